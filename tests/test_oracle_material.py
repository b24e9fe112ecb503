"""CPU tests pinning the oracle's per-voxel stress update (drive_eps_sig.f,
polar.f, mm10_a.f, mm10_b.f, cep2A.f) by mathematics the update must obey.

Pins: closed-form rotations (Kocks matrix, orthogonality, round trip), the
polar decomposition against an SVD (within the reference Cardano method's
own rounding noise), the elastic limit sigma = C:d of the explicit update,
finite differences of the residual (local Jacobian) and of P(F) (the exact
dP/dF of cep2A), the discrete Voce law at the converged state, and
rotation of the plastic spin.  The reference ships no golden data for any
of these.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
from cpfft_amd.microstructure import kocks_angles, kocks_matrix, random_rotations

ANG = np.array([30.0, 40.0, 50.0])
I3 = np.eye(3)


def upd(c, Fn, Fn1, sn, step=1, it=1, dt=0.5, ang=ANG):
    return O.voxel_update(c, ang, step, it, dt, Fn, Fn1, sn)


def test_rotation_matrix_kocks():
    rng = np.random.default_rng(0)
    R = random_rotations(rng, 20)
    a = kocks_angles(R)
    for k in range(20):
        g = np.zeros(9)
        O.lib().orc_rotation_matrix(O.P(np.ascontiguousarray(a[k])), O.P(g))
        g = g.reshape(3, 3)
        assert np.allclose(g, R[k], atol=1e-13)
        assert np.allclose(g @ g.T, I3, atol=1e-14)
        assert abs(np.linalg.det(g) - 1) < 1e-14
    assert np.allclose(kocks_matrix(a), R, atol=1e-13)


@pytest.mark.parametrize("amp", [1e-2, 1e-4, 1e-6])
def test_polar_decomposition_vs_svd(amp):
    """R = F U^-1 (rtcmp1) agrees with the SVD polar factor; the Cardano
    eigenvalue step of the reference loses up to ~5e-9 near 1e-3 strain."""
    rng = np.random.default_rng(1)
    worst = 0.0
    for _ in range(50):
        F = I3 + amp * rng.standard_normal((3, 3))
        R = np.zeros(9)
        O.lib().orc_rtcmp1(O.P(np.ascontiguousarray(F.ravel())), O.P(R))
        U, s, Vt = np.linalg.svd(F)
        worst = max(worst, abs(R.reshape(3, 3) - U @ Vt).max())
    assert worst < 1e-8


def test_zero_load_is_zero():
    c = O.crystal()
    s0 = O.state_init(c, 1)[:, 0]
    P, K, s1, st, fail = upd(c, I3, I3, s0, step=1, it=0)
    assert fail == 0 and not P.any()
    assert s1[15] == c.tau_y + 1e-5


def test_elastic_predictor_is_hypoelastic():
    """Iteration 0 from a stress-free state: sigma_unrot = C:d exactly
    (mm10_a.f:2735-2753 with sigma_n = 0), d = Q(R_{n+1/2}) sym(dF Fh^-1)."""
    c = O.crystal()
    s0 = O.state_init(c, 1)[:, 0]
    rng = np.random.default_rng(2)
    Fn1 = I3 + 1e-4 * rng.standard_normal((3, 3))
    P, K, s1, st, fail = upd(c, I3, Fn1, s0, step=1, it=0)
    g = kocks_matrix(ANG)
    # crystal stiffness in Voigt form (engineering shear)
    Cc = np.array(c.stiff[:]).reshape(6, 6)
    Fh = 0.5 * (I3 + Fn1)
    U, sv, Vt = np.linalg.svd(Fh)
    Rh = U @ Vt
    L = (Fn1 - I3) @ np.linalg.inv(Fh)
    d = Rh.T @ (0.5 * (L + L.T)) @ Rh
    # fourth-order crystal stiffness from its Voigt matrix, rotated to the
    # sample frame index by index: C'_ijkl = t_ia t_jb t_kc t_ld C_abcd
    vi = [[0, 3, 5], [3, 1, 4], [5, 4, 2]]
    C4 = np.array([[[[Cc[vi[a][b], vi[c][e]] for e in range(3)] for c in range(3)]
                    for b in range(3)] for a in range(3)])
    t = g.T
    Cr = np.einsum("ia,jb,kc,ld,abcd->ijkl", t, t, t, t, C4)
    sig = np.einsum("ijkl,kl->ij", Cr, d)
    sv6 = np.array([sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[1, 2], sig[0, 2]])
    assert np.allclose(s1[0:6], sv6, rtol=1e-7, atol=1e-10)


def test_cep2A_is_exact_derivative_of_elastic_update():
    """cep2A.f:9-10 claims dP/dF equals the finite difference of the update;
    at iteration 0 from a stress-free state the update is exactly
    differentiable and K4 matches a central difference of P(F_{n+1})."""
    c = O.crystal()
    s0 = O.state_init(c, 1)[:, 0]
    rng = np.random.default_rng(3)
    F1 = I3 + 1e-4 * rng.standard_normal((3, 3))
    P, K, _, _, _ = upd(c, I3, F1, s0, it=0)
    h = 1e-7
    Kfd = np.zeros((9, 9))
    for m in range(9):
        d = np.zeros(9)
        d[m] = h
        Kfd[:, m] = (upd(c, I3, F1 + d.reshape(3, 3), s0, it=0)[0] -
                     upd(c, I3, F1 - d.reshape(3, 3), s0, it=0)[0]) / (2 * h)
    assert abs(K.reshape(9, 9) - Kfd).max() <= 1e-7 * abs(Kfd).max()


def test_consistent_tangent_plastic():
    """Plastic update at tight local tolerances: K4 = dP/dF up to the
    symmetrisation of the material tangent (mm10_a.f:1141-1142)."""
    c = O.crystal(jac_exact=1)
    c.rtol = c.atol = c.atol1 = c.rtol1 = 1e-12
    s0 = O.state_init(c, 1)[:, 0]
    rng = np.random.default_rng(3)
    F1 = np.diag([1 + 1e-3, 1 - 5e-4, 1 - 5e-4]) + 1e-4 * rng.standard_normal((3, 3))
    P, K, s1, st, fail = upd(c, I3, F1, s0)
    assert fail == 0 and st[1] > 0
    h = 1e-6
    Kfd = np.zeros((9, 9))
    for m in range(9):
        d = np.zeros(9)
        d[m] = h
        Kfd[:, m] = (upd(c, I3, F1 + d.reshape(3, 3), s0)[0] -
                     upd(c, I3, F1 - d.reshape(3, 3), s0)[0]) / (2 * h)
    assert abs(K.reshape(9, 9) - Kfd).max() <= 2e-3 * abs(Kfd).max()


def _local_RJ(c, Fn1, x):
    L = O.lib()
    L.orc_local_RJ.argtypes = [ctypes.POINTER(O.OrcCrystal), O._dp, ctypes.c_double, O._dp,
                               O._dp, O._dp, O._dp, O._dp, O._dp]
    s0 = O.state_init(c, 1)[:, 0]
    R = np.zeros(7)
    J = np.zeros(49)
    L.orc_local_RJ(ctypes.byref(c), O.P(ANG), 0.5, O.P(np.eye(3).ravel().copy()),
                   O.P(np.ascontiguousarray(Fn1.ravel())), O.P(s0), O.P(np.ascontiguousarray(x)),
                   O.P(R), O.P(J))
    return R, J.reshape(7, 7)


@pytest.mark.parametrize("jac_exact", [0, 1])
def test_local_jacobian_vs_finite_difference(jac_exact):
    """J11, J12, J21 are the derivatives of R (mm10_b.f:177-481, 1918-1949).
    J22 as mm10_ehard_voche writes it carries a spurious sign(slipinc)
    factor; the exact mode drops it and matches the difference quotient."""
    c = O.crystal(jac_exact=jac_exact)
    F1 = np.diag([1 + 1e-3, 1 - 5e-4, 1 - 5e-4])
    x = np.array([33.32, -11.46, -22.16, 1.0, -2.0, 0.5, 16.3])
    R, J = _local_RJ(c, F1, x)
    Jfd = np.zeros((7, 7))
    for m in range(7):
        h = 1e-6 * max(1.0, abs(x[m]))
        xp, xm = x.copy(), x.copy()
        xp[m] += h
        xm[m] -= h
        Jfd[:, m] = (_local_RJ(c, F1, xp)[0] - _local_RJ(c, F1, xm)[0]) / (2 * h)
    err = abs(J - Jfd) / (abs(Jfd) + 1e-8 * abs(Jfd).max())
    assert err[:, :6].max() < 1e-6 and err[:6, 6].max() < 1e-6
    if jac_exact:
        assert err[6, 6] < 1e-6
    else:
        assert err[6, 6] > 0.1


def test_voce_discrete_law_at_converged_state():
    """At convergence R2 = 0: tau~ = tau~_n + theta_0 (1 - (tau~-tau_y)/tau_v) S
    with S = sum |slip increments| (mm10_b.f:1888-1916, voce_m = 1)."""
    c = O.crystal(jac_exact=1)
    c.rtol = c.atol = c.atol1 = c.rtol1 = 1e-12
    s0 = O.state_init(c, 1)[:, 0]
    F1 = np.diag([1 + 1e-3, 1 - 5e-4, 1 - 5e-4])
    P, K, s1, st, fail = upd(c, I3, F1, s0)
    tt = s1[15]
    S = abs(s1[23:35]).sum()
    tt_n = c.tau_y + 1e-5
    closed = (tt_n + c.theta_0 * S * (1 + c.tau_y / c.tau_v)) / (1 + c.theta_0 * S / c.tau_v)
    assert abs(tt - closed) <= 1e-9 * tt
    assert tt > tt_n


def test_plastic_rotation_is_orthogonal_and_small():
    c = O.crystal(jac_exact=1)
    s0 = O.state_init(c, 1)[:, 0]
    F1 = np.diag([1 + 1e-3, 1 - 5e-4, 1 - 5e-4])
    P, K, s1, st, fail = upd(c, I3, F1, s0)
    Rp = s1[6:15].reshape(3, 3).T
    assert np.allclose(Rp @ Rp.T, I3, atol=1e-14)
    assert 0 < abs(Rp - I3).max() < 1e-2


def test_local_failure_reverts_to_state_n():
    """A strain increment far outside the local solver's reach: up to four
    step halvings (mm10_a.f:2661-2800), then the voxel keeps state n."""
    c = O.crystal(jac_exact=0)
    s0 = O.state_init(c, 1)[:, 0]
    F1 = np.diag([1.02, 0.99, 0.99])
    P, K, s1, st, fail = upd(c, I3, F1, s0)
    assert fail == 1 and st[2] == 5
    assert np.array_equal(s1[0:6], s0[0:6]) and s1[15] == s0[15]


def test_P_is_first_piola_of_cauchy():
    """P = J sigma F^-T with sigma = R t R^T (drive_eps_sig.f:262-273)."""
    c = O.crystal(jac_exact=1)
    s0 = O.state_init(c, 1)[:, 0]
    rng = np.random.default_rng(9)
    F1 = np.diag([1 + 8e-4, 1 - 4e-4, 1 - 4e-4]) + 2e-4 * rng.standard_normal((3, 3))
    P, K, s1, st, fail = upd(c, I3, F1, s0)
    t = s1[0:6]
    T = np.array([[t[0], t[3], t[5]], [t[3], t[1], t[4]], [t[5], t[4], t[2]]])
    U, sv, Vt = np.linalg.svd(F1)
    R = U @ Vt
    sig = R @ T @ R.T
    Pe = np.linalg.det(F1) * sig @ np.linalg.inv(F1).T
    assert np.allclose(P.reshape(3, 3), Pe, rtol=1e-7, atol=1e-8 * abs(Pe).max())


def test_voronoi_kdtree_matches_brute_force():
    from cpfft_amd.microstructure import voronoi
    a = voronoi(np.random.default_rng(4), 24, 100, method="brute")
    b = voronoi(np.random.default_rng(4), 24, 100, method="kdtree")
    assert np.array_equal(a, b)
