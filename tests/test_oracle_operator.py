"""CPU tests pinning the oracle's projection operator, CG, NBC_update and
tangent_homo (G_K_dF.f, FFT_nr3.f, tangent_homo.f).

Pins: an independent numpy restatement of the literal complex-to-complex +
phase-shift path; the closed-form real multiplier M(f) of the reference's
even-N operator; projection identities (idempotence, zero mean) of the
textbook operator; dense linear algebra for NBC_update; homogeneous media
for tangent_homo.  No reference test holds golden data for this path.
"""
import numpy as np
import pytest

import oracle_lib as O


def numpy_literal(F, N, shift):
    """G_K_dF.f:11-227 restated with numpy.fft: phase e^{-i theta} with
    theta = 2 pi shift (i+j+k)/N, fftn, Ghat4 on q = idx - Nhalf, ifftn, real
    part of the inverse phase."""
    Nhalf = (N + 1) // 2 if N % 2 else N // 2 + 1
    idx = np.arange(N)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    ph = np.exp(-2j * np.pi * shift * (I + J + K) / N)
    q = np.stack([I + 1 - Nhalf, J + 1 - Nhalf, K + 1 - Nhalf]).astype(float)
    qq = (q ** 2).sum(0)
    zero = qq < 1e-10
    qq = np.where(zero, 1.0, qq)
    Fh = np.fft.fftn(F.reshape(3, 3, N, N, N) * ph, axes=(2, 3, 4))
    Aq = np.einsum("ilxyz,lxyz->ixyz", Fh, q)
    G = np.einsum("ixyz,jxyz->ijxyz", Aq, q) / qq
    G = np.where(zero, 0.0, G)
    return np.real(np.fft.ifftn(G, axes=(2, 3, 4)) * np.conj(ph)).reshape(9, -1), q


def wrap(v, N):
    return ((v + N // 2) % N) - N // 2


def r2c_multiplier(F, N, mode):
    """Closed form of the operator on the real half spectrum:
    reference, even N: M(f) = (Gamma(w(f-1)) + Gamma(w(-f-1)))/2;
    textbook: Gamma(w(f)) with the Nyquist planes removed."""
    Fh = np.fft.rfftn(F.reshape(3, 3, N, N, N), axes=(2, 3, 4))
    fx = np.fft.fftfreq(N, 1.0 / N)
    fz = np.arange(N // 2 + 1)
    f = np.stack(np.meshgrid(fx, fx, fz, indexing="ij")).astype(float)

    def gam(qv, w):
        qq = (qv ** 2).sum(0)
        m = qq < 1e-10
        qq = np.where(m, 1.0, qq)
        Aq = np.einsum("ilxyz,lxyz->ixyz", Fh, qv)
        return w * np.where(m, 0.0, np.einsum("ixyz,jxyz->ijxyz", Aq, qv) / qq)

    if N % 2 == 1:
        out = gam(wrap(f, N), 1.0)
    elif mode == 0:
        out = gam(wrap(f - 1, N), 0.5) + gam(wrap(-f - 1, N), 0.5)
    else:
        out = gam(wrap(f, N), 1.0)
        out[:, :, (np.abs(f) == N // 2).any(0)] = 0.0
    return np.fft.irfftn(out, s=(N, N, N), axes=(2, 3, 4)).reshape(9, -1)


@pytest.mark.parametrize("N", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("mode", [0, 1])
def test_gkdf_matches_numpy_restatements(N, mode):
    rng = np.random.default_rng(100 + N)
    F = rng.standard_normal((9, N ** 3))
    K4 = rng.standard_normal((81, N ** 3))
    fo = O.FFT(N, mode)
    Nhalf = (N + 1) // 2 if N % 2 else N // 2 + 1
    shift = N // 2 if (mode == 1 and N % 2 == 0) else Nhalf
    a = fo.gkdf(F)
    lit, _ = numpy_literal(F, N, shift)
    if mode == 1 and N % 2 == 0:
        # the literal path with the corrected shift still needs the Nyquist cut
        lit = r2c_multiplier(F, N, 1)
    scale = abs(a).max()
    assert abs(a - lit).max() <= 1e-13 * scale
    assert abs(a - r2c_multiplier(F, N, mode)).max() <= 1e-13 * scale
    # with K4: ddot42n first
    KF = np.einsum("rcv,cv->rv", K4.reshape(9, 9, -1), F)
    b = fo.gkdf(F, K4)
    assert abs(b - r2c_multiplier(KF, N, mode)).max() <= 1e-12 * abs(b).max()


def test_ddot42n_order():
    """ddot42n's pairwise order t1 + (((t2+t6)+(t4+t8)) + ((t3+t7)+(t5+t9)))
    (G_K_dF.f:258-265): pinned on inputs where the order matters."""
    V = 1
    A = np.zeros((81, V))
    B = np.zeros((9, V))
    A[:9, 0] = 1.0
    B[:, 0] = [0.0, 1.0, 0.0, 1e16, 0.0, 1.0, 0.0, -1e16, 0.0]
    C = np.zeros((9, V))
    O.lib().orc_ddot42n(V, O.P(A), O.P(B), O.P(C))
    t = B[:, 0]
    expect = t[0] + (((t[1] + t[5]) + (t[3] + t[7])) + ((t[2] + t[6]) + (t[4] + t[8])))
    assert C[0, 0] == expect
    assert expect != sum(t)  # a plain left-to-right sum differs


@pytest.mark.parametrize("N", [4, 6, 8])
def test_reference_operator_even_N_quirk(N):
    """For even N the reference operator is not a projection and does not
    annihilate the mean: mean(G:A)_ij = 1/3 sum_l mean(A)_il."""
    rng = np.random.default_rng(N)
    A = rng.standard_normal((9, N ** 3))
    fo = O.FFT(N, 0)
    G = fo.gkdf(A)
    mA = A.mean(1).reshape(3, 3)
    expect = np.repeat(mA.sum(1, keepdims=True) / 3.0, 3, axis=1).ravel()
    assert np.allclose(G.mean(1), expect, atol=1e-13)
    assert abs(fo.gkdf(G) - G).max() > 1e-3  # not idempotent


@pytest.mark.parametrize("N", [5, 6, 7, 8])
def test_textbook_operator_is_projection(N):
    rng = np.random.default_rng(N)
    A = rng.standard_normal((9, N ** 3))
    fo = O.FFT(N, 1)
    G = fo.gkdf(A)
    assert abs(fo.gkdf(G) - G).max() <= 1e-13 * abs(G).max()
    assert abs(G.mean(1)).max() <= 1e-14
    # compatible fields: Gamma of a gradient field returns it (curl-free part)
    u = rng.standard_normal((3, N, N, N))
    uh = np.fft.fftn(u, axes=(1, 2, 3))
    fx = np.fft.fftfreq(N, 1.0 / N)
    k = np.stack(np.meshgrid(fx, fx, fx, indexing="ij"))
    if N % 2 == 0:
        uh[:, (np.abs(k) == N // 2).any(0)] = 0.0
    grad = np.real(np.fft.ifftn(1j * uh[:, None] * k[None, :], axes=(2, 3, 4)))
    H = np.ascontiguousarray(grad.reshape(9, -1))
    assert abs(fo.gkdf(H) - H).max() <= 1e-12 * abs(H).max()


def test_odd_N_modes_agree():
    rng = np.random.default_rng(3)
    F = rng.standard_normal((9, 7 ** 3))
    assert np.array_equal(O.FFT(7, 0).gkdf(F), O.FFT(7, 1).gkdf(F))


def _elastic_K4(c, N, seed=11):
    from cpfft_amd.microstructure import polycrystal
    _, ang = polycrystal(N, 4, seed)
    V = N ** 3
    I = np.zeros((9, V))
    I[[0, 4, 8]] = 1.0
    P, K4, _, _ = O.drive_eps_sig(c, ang, 1, 0, 0.5, I, I, O.state_init(c, V))
    return K4


@pytest.mark.parametrize("N,mode", [(5, 0), (7, 1), (8, 1), (6, 0)])
def test_pcg_solves_system(N, mode):
    """fftPcg: x with G(K4:x) = b to the stop test, iteration count below the
    1000 cap; x stays in the compatible subspace for the projection."""
    c = O.crystal()
    K4 = _elastic_K4(c, N)
    fo = O.FFT(N, mode)
    dF = np.zeros((9, N ** 3))
    dF[0] = 1e-3
    b = -fo.gkdf(dF, K4)
    x, rc, it, rel = fo.pcg(K4, b, 1e-10)
    assert rc == 0 and 0 < it < 1000
    r = b - fo.gkdf(x, K4)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.0001
    if mode == 1 or N % 2 == 1:
        assert abs(fo.gkdf(x) - x).max() <= 1e-10 * abs(x).max()


def test_pcg_edge_cases():
    c = O.crystal()
    K4 = _elastic_K4(c, 5)
    fo = O.FFT(5, 0)
    z = np.zeros((9, 125))
    x, rc, it, rel = fo.pcg(K4, z, 1e-8)
    assert rc == 1 and not x.any()   # zero right-hand side (FFT_nr3.f:256-260)
    assert fo.pcg(K4, z + 1.0, 0.0)[1] == -1   # improper tolerance (FFT_nr3.f:248)
    assert fo.pcg(K4, z + 1.0, 1.0)[1] == -1


def test_nbc_update_dense():
    rng = np.random.default_rng(5)
    C = rng.standard_normal((9, 9)) + 10 * np.eye(9)
    isNBC = np.array([0, 0, 0, 0, 1, 0, 0, 0, 1], np.int32)
    DbarF = rng.standard_normal(9)
    Pbar = rng.standard_normal(9)
    PBC = rng.standard_normal(9)
    d, rc = O.nbc_update(C.ravel(), DbarF, Pbar, PBC, isNBC)
    A = np.eye(9)
    bb = DbarF.copy()
    for i in np.nonzero(isNBC)[0]:
        A[i] = C.ravel()[9 * i:9 * i + 9]
        bb[i] = PBC[i] - Pbar[i]
    assert rc == 0 and np.allclose(d, np.linalg.solve(A, bb), rtol=1e-13, atol=1e-15)
    # all six shears prescribed as stresses at a skew-insensitive tangent: singular
    Cs = np.zeros(81)
    _, rc = O.nbc_update(Cs, DbarF, Pbar, PBC, np.ones(9, np.int32))
    assert rc != 0


def test_tangent_homo_homogeneous_odd_N():
    """Homogeneous medium: the fluctuation vanishes and C_homo = K4 exactly
    (tangent_homo.f:42-68 with Gamma(0) = 0)."""
    c = O.crystal()
    N = 5
    V = N ** 3
    ang = np.tile(np.array([[20.0], [35.0], [70.0]]), (1, V))
    I = np.zeros((9, V))
    I[[0, 4, 8]] = 1.0
    _, K4, _, _ = O.drive_eps_sig(c, ang, 1, 0, 0.5, I, I, O.state_init(c, V))
    Ch = O.FFT(N, 0).tangent_homo(K4, 1e-8)
    K = K4[:, 0].reshape(9, 9)
    # C_homo[9(j-1)+i] = dPbar_j / dFbar_i = K[j, i]
    assert np.allclose(Ch.reshape(9, 9), K, rtol=1e-13, atol=1e-9)
