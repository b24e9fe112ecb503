"""CPU check of the library's per-voxel update algebra: the device source
(cpfft_amd/csrc/mm10_core.h) compiled for the host (tests/hostcheck) against
the oracle's drive_eps_sig on identical inputs.

The device code accumulates every sum over slip systems in the lattice frame
and rotates it once; the oracle follows the reference and rotates each slip
system's Schmid tensor.  Same mathematics, different rounding: per-field
agreement within 1e-9 relative (the bar of the GPU parity tests).
"""
import ctypes
import os

import numpy as np
import pytest

import cpfft_amd as C
import oracle_lib as O
from cpfft_amd.microstructure import polycrystal

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "hostcheck", "libvoxel_host.so")


def host_lib():
    if not os.path.exists(LIB):
        pytest.skip("host check library not built (python -m cpfft_amd.build hostcheck)")
    L = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.cpfft_host_drive_eps_sig.argtypes = [ctypes.POINTER(C.CMaterial), ctypes.c_double,
                                           ctypes.c_long, dp, ctypes.c_int, ctypes.c_int, dp, dp,
                                           dp, dp, dp, dp, ctypes.POINTER(ctypes.c_longlong)]
    return L


def P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def host_drive(mat, dt, ang, step, it, Fn, Fn1, sn):
    L = host_lib()
    cm = C.CMaterial(**{k: getattr(mat, k) for k, _ in C.CMaterial._fields_})
    V = Fn.shape[1]
    snp1 = np.zeros_like(sn)
    Pn1 = np.zeros((9, V))
    K4 = np.zeros((81, V))
    cnt = (ctypes.c_longlong * 6)()
    L.cpfft_host_drive_eps_sig(ctypes.byref(cm), dt, V, P(np.ascontiguousarray(ang)), step, it,
                               P(np.ascontiguousarray(Fn)), P(np.ascontiguousarray(Fn1)),
                               P(np.ascontiguousarray(sn)), P(snp1), P(Pn1), P(K4), cnt)
    return Pn1, K4, snp1, list(cnt)


def rel(a, b):
    return abs(a - b).max() / max(abs(b).max(), 1e-300)


def near_identity(rng, V, amp, uniaxial=0.0):
    F = np.zeros((9, V))
    F[[0, 4, 8]] = 1.0
    F += amp * rng.standard_normal((9, V))
    F[0] += uniaxial
    F[4] -= 0.5 * uniaxial
    F[8] -= 0.5 * uniaxial
    return F


def check(a, b, nslip, what):
    Pn, Kn, sn, cn = a
    Po, Ko, so, st = b
    assert cn[3] == st[3], what
    assert rel(Pn, Po) < 1e-9, what
    assert rel(Kn, Ko) < 1e-9, what
    for w0, w1 in [(0, 6), (6, 15), (15, 17), (17, 23), (23, 23 + 2 * nslip),
                   (23 + 2 * nslip, 26 + 2 * nslip)]:
        assert rel(sn[w0:w1], so[w0:w1]) < 1e-9, (what, w0)


# seed offset per slip set: 1 fcc, 80 bcc24 (numbered 8 when its seeds were
# chosen), 8 bcc48
SEED_OF = {1: 1, 80: 8, 8: 48, 2: 2, 6: 6, 7: 7}


@pytest.mark.parametrize("slip_set,jac_exact",
                         [(1, 0), (1, 1), (80, 1), (8, 0), (2, 1), (6, 0), (7, 1)])
def test_host_update_matches_oracle(slip_set, jac_exact):
    V = 6 ** 3
    rng = np.random.default_rng(40 + SEED_OF[slip_set] + jac_exact)
    _, ang = polycrystal(6, 5, 11 + SEED_OF[slip_set])
    mat = C.Material(slip_set=slip_set, jac_exact=jac_exact)
    oc = O.crystal(slip_set=slip_set, jac_exact=jac_exact)
    dt = 0.25
    Fn = near_identity(rng, V, 0.0)
    Fn1 = near_identity(rng, V, 1e-4, uniaxial=6e-4)
    sn = O.state_init(oc, V)
    for step, it in [(1, 0), (1, 1), (1, 2)]:
        h = host_drive(mat, dt, ang, step, it, Fn, Fn1, sn)
        o = O.drive_eps_sig(oc, ang, step, it, dt, Fn, Fn1, sn)
        check(h, o, oc.nslip, (step, it))
    # a second load step from the converged state
    sn = o[2]
    Fn = Fn1
    Fn1 = Fn + near_identity(rng, V, 5e-5, uniaxial=5e-4) - near_identity(rng, V, 0.0)
    for it in (1, 2):
        h = host_drive(mat, dt, ang, 2, it, Fn, Fn1, sn)
        o = O.drive_eps_sig(oc, ang, 2, it, dt, Fn, Fn1, sn)
        check(h, o, oc.nslip, (2, it))
    assert h[3][1] > 0  # plastic Newton iterations happened


def test_warm_start_from_exact_slips_cuts_local_iterations():
    """The warm-start mechanism (mm10_core.h warm_x0 and the predictor's
    choice of start) with a perfect prediction: started from the converged
    slip increments of the same update, the local solves converge to the
    same state (the Newton tolerances are unchanged) in far fewer
    iterations than from the reference's start."""
    L = host_lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.cpfft_host_drive_eps_sig_ws.argtypes = L.cpfft_host_drive_eps_sig.argtypes + [
        ctypes.POINTER(ctypes.c_float)]
    V = 6 ** 3
    rng = np.random.default_rng(3)
    _, ang = polycrystal(6, 5, 12)
    mat = C.Material(jac_exact=1)
    oc = O.crystal(jac_exact=1)
    dt = 0.25
    Fn = near_identity(rng, V, 0.0)
    Fn1 = near_identity(rng, V, 1e-4, uniaxial=6e-4)
    _, _, s1, _ = host_drive(mat, dt, ang, 1, 1, Fn, Fn1, O.state_init(oc, V))
    Fn, Fn1 = Fn1, Fn1 + near_identity(rng, V, 2e-5, uniaxial=2.5e-4) - near_identity(rng, V, 0.0)
    Pp, Kp, sp, cp = host_drive(mat, dt, ang, 2, 1, Fn, Fn1, s1)
    g = sp[23:35]
    u = np.sign(g) * (np.abs(g) / (mat.gamma_bar * dt)) ** (1.0 / mat.rate_n)
    ws = np.ascontiguousarray(u.astype(np.float32))
    cm = C.CMaterial(**{k: getattr(mat, k) for k, _ in C.CMaterial._fields_})
    snp1 = np.zeros_like(s1)
    Pn1 = np.zeros((9, V))
    K4 = np.zeros((81, V))
    cnt = (ctypes.c_longlong * 6)()
    L.cpfft_host_drive_eps_sig_ws(ctypes.byref(cm), dt, V, P(np.ascontiguousarray(ang)), 2, 1,
                                  P(Fn), P(np.ascontiguousarray(Fn1)), P(s1), P(snp1), P(Pn1),
                                  P(K4), cnt, ws.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    plain = cp[0] + cp[1]
    warm = cnt[0] + cnt[1]
    assert cnt[5] == 0
    assert warm < 0.5 * plain, (warm, plain)
    # both solves stop within the local tolerances (atol 1e-5 MPa, rtol
    # 5e-5 of the starting residual), at different points of that ball
    assert rel(Pn1, Pp) < 1e-6
    assert rel(snp1[0:6], sp[0:6]) < 1e-6


@pytest.mark.parametrize("slip_set", [1, 80, 8])
def test_warm_start_from_previous_iterate_cuts_local_iterations(slip_set):
    """The warm start from the step's previous global iterate (the state
    n+1 the update overwrites): for a global Newton correction of the size
    the later iterations make, the local solves reach the same state (the
    tolerances are unchanged) in far fewer iterations than from the
    reference's start, for every slip set."""
    L = host_lib()
    L.cpfft_host_drive_eps_sig_prev.argtypes = L.cpfft_host_drive_eps_sig.argtypes
    V = 5 ** 3
    rng = np.random.default_rng(11)
    _, ang = polycrystal(5, 4, 13)
    mat = C.Material(slip_set=slip_set, jac_exact=1)
    oc = O.crystal(slip_set=slip_set, jac_exact=1)
    dt = 0.25
    Fn = near_identity(rng, V, 0.0)
    Fn1 = near_identity(rng, V, 1e-4, uniaxial=6e-4)
    _, _, s1, _ = host_drive(mat, dt, ang, 1, 1, Fn, Fn1, O.state_init(oc, V))
    Fn, Fa = Fn1, Fn1 + near_identity(rng, V, 2e-5, uniaxial=2.5e-4) - near_identity(rng, V, 0.0)
    _, _, sa, ca = host_drive(mat, dt, ang, 2, 1, Fn, Fa, s1)
    assert ca[0] + ca[1] > 0
    Fb = Fa + near_identity(rng, V, 1e-7) - near_identity(rng, V, 0.0)
    Pp, Kp, sp, cp = host_drive(mat, dt, ang, 2, 2, Fn, Fb, s1)
    cm = C.CMaterial(**{k: getattr(mat, k) for k, _ in C.CMaterial._fields_})
    snp1 = sa.copy()
    Pn1 = np.zeros((9, V))
    K4 = np.zeros((81, V))
    cnt = (ctypes.c_longlong * 6)()
    assert L.cpfft_host_drive_eps_sig_prev(ctypes.byref(cm), dt, V, P(np.ascontiguousarray(ang)),
                                           2, 2, P(Fn), P(np.ascontiguousarray(Fb)), P(s1),
                                           P(snp1), P(Pn1), P(K4), cnt) == 0
    plain, warm = cp[0] + cp[1], cnt[0] + cnt[1]
    print(f"\nslip set {slip_set}: local iterations plain {plain} warm {warm}; "
          f"P rel diff {rel(Pn1, Pp):.2e}")
    assert cnt[3] == 0 and cnt[5] == 0
    assert warm < 0.6 * plain, (warm, plain)
    assert rel(Pn1, Pp) < 1e-6
    assert rel(snp1[0:6], sp[0:6]) < 1e-6
    # the consistent tangent is formed at the second-to-last local iterate
    # (mm10_a.f:1134-1142), which differs between the two solves
    assert rel(K4, Kp) < 1e-3


def rotations(rng, V):
    """V random rotation matrices as [9][V] (row-major per voxel)"""
    q = rng.standard_normal((V, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)])
    return R


def matmul9(A, B):
    return np.einsum("ikv,kjv->ijv", A.reshape(3, 3, -1), B.reshape(3, 3, -1)).reshape(9, -1)


@pytest.mark.parametrize("slip_set", [1, 80])
def test_network_inputs_match_fp64(slip_set):
    """The warm-start network's inputs (mm10_core.h voxel_features, FP32:
    rotation of F_n+1/2 by Higham's iteration, FP32 lattice rotations and
    slip_to_u) against the same quantities in FP64 (step_kinematics's polar
    rotation, FP64 slip_to_u), at a first step, at a plastic second step,
    and under finite rotations and 30% stretches.  The network reads them in
    BF16 (relative rounding 2^-9), so 1e-5 of each block's scale is far
    below what it can see."""
    L = host_lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.cpfft_host_features.argtypes = [ctypes.POINTER(C.CMaterial), ctypes.c_double, ctypes.c_long,
                                      dp, ctypes.c_int, dp, dp, dp,
                                      ctypes.POINTER(ctypes.c_float), dp]
    V = 6 ** 3
    rng = np.random.default_rng(70 + slip_set)
    _, ang = polycrystal(6, 5, 21 + slip_set)
    mat = C.Material(slip_set=slip_set, jac_exact=1)
    oc = O.crystal(slip_set=slip_set, jac_exact=1)
    ns = oc.nslip
    nin = 18 + 2 * ns
    cm = C.CMaterial(**{k: getattr(mat, k) for k, _ in C.CMaterial._fields_})
    dt = 0.25

    def feats(step, Fn, Fn1, sn):
        f32 = np.zeros((V, nin), np.float32)
        f64 = np.zeros((V, nin))
        rc = L.cpfft_host_features(ctypes.byref(cm), dt, V, P(np.ascontiguousarray(ang)), step,
                                   P(np.ascontiguousarray(Fn)), P(np.ascontiguousarray(Fn1)),
                                   P(np.ascontiguousarray(sn)),
                                   f32.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), P(f64))
        assert rc == 0
        return f32.astype(np.float64), f64

    def blocks_agree(f32, f64, what):
        for lo, hi, name in [(0, 9, "stress"), (9, 9 + ns, "tt"), (9 + ns, 18 + ns, "D"),
                             (18 + ns, nin, "u")]:
            scale = abs(f64[:, lo:hi]).max()
            if scale == 0.0:
                assert abs(f32[:, lo:hi]).max() == 0.0, (what, name)
                continue
            err = abs(f32[:, lo:hi] - f64[:, lo:hi]).max() / scale
            assert err < 1e-5, (what, name, err)

    Fn = near_identity(rng, V, 0.0)
    Fn1 = near_identity(rng, V, 1e-4, uniaxial=6e-4)
    sn = O.state_init(oc, V)
    blocks_agree(*feats(1, Fn, Fn1, sn), "step 1")
    _, _, s1, cnt = host_drive(mat, dt, ang, 1, 1, Fn, Fn1, sn)
    Fn, Fn1 = Fn1, Fn1 + near_identity(rng, V, 2e-5, uniaxial=2.5e-4) - near_identity(rng, V, 0.0)
    _, _, s2, cnt = host_drive(mat, dt, ang, 2, 1, Fn, Fn1, s1)
    assert cnt[1] > 0  # the state carries slip
    f32, f64 = feats(3, Fn1, Fn1 + near_identity(rng, V, 2e-5, uniaxial=2.5e-4)
                     - near_identity(rng, V, 0.0), s2)
    assert abs(f64[:, 18 + ns:]).max() > 0.1  # non-trivial u of the slipping systems
    blocks_agree(f32, f64, "plastic step")
    # finite rotation and stretch: F = R (I + 0.3 sym noise + stretch)
    R = rotations(rng, V)
    U = near_identity(rng, V, 0.05, uniaxial=0.3)
    U = 0.5 * (U + U.reshape(3, 3, V).transpose(1, 0, 2).reshape(9, V))
    Fa = matmul9(R, U)
    Fb = matmul9(R, U + near_identity(rng, V, 1e-4, uniaxial=5e-4) - near_identity(rng, V, 0.0))
    blocks_agree(*feats(3, Fa, Fb, s2), "rotated and stretched")
