"""GPU parity of the displacement reconstruction (cpfft_f2disp, the
reference's f2disp.f) against the oracle's sparse direct solve of the same
brick-mesh system, and at full grid size through properties fixed by the
mathematics: a homogeneous F gives u = (F - I) X exactly, and u is linear
in F - I.  Both solvers are exact up to rounding of the positions x = X + u,
whose size is the box edge: tolerances are 1e-10 of the longest edge."""
import numpy as np
import pytest

import cpfft_amd as C
from cpfft_amd.microstructure import polycrystal
from oracle import f2disp_oracle as O

pytestmark = pytest.mark.gpu


def _solver(N):
    _, ang = polycrystal(N, 2, 9)
    return C.Solver(N, ang, C.Material(), dt=0.25)


@pytest.mark.parametrize("N,L", [(4, (1.0, 1.0, 1.0)), (7, (2.0, 1.0, 0.5)), (16, (1.0, 3.0, 2.0))])
def test_f2disp_against_oracle(N, L):
    rng = np.random.default_rng(100 + N)
    F = np.eye(3).reshape(9, 1) + 0.05 * rng.standard_normal((9, N ** 3))
    s = _solver(N)
    s.upload(C.FN1, F)
    u = s.f2disp(C.FN1, L)
    ref = O.f2disp(F, N, L)
    err = np.abs(u - ref).max()
    print(f"\n[f2disp] N={N} L={L} max|u|={np.abs(ref).max():.3e} max err={err:.3e}", flush=True)
    assert u.shape == ref.shape
    assert np.all(u[0] == 0.0)
    assert err <= 1e-10 * max(L)
    s.close()


@pytest.mark.parametrize("N", [64, 128])
def test_f2disp_full_size_properties(N):
    rng = np.random.default_rng(7 * N)
    L = (1.0, 1.5, 0.75)
    s = _solver(N)
    X = O.node_coords(N, L)
    # homogeneous F: trilinear bricks reproduce the affine map exactly
    G = 0.02 * rng.standard_normal((3, 3))
    s.upload(C.FN1, np.tile((np.eye(3) + G).reshape(9, 1), (1, N ** 3)))
    u = s.f2disp(C.FN1, L)
    err = np.abs(u - X @ G.T).max()
    print(f"\n[f2disp] N={N} homogeneous max err={err:.3e}", flush=True)
    assert err <= 1e-10 * max(L)
    # linearity in F - I
    D1 = 0.01 * rng.standard_normal((9, N ** 3))
    D2 = 0.01 * rng.standard_normal((9, N ** 3))
    I9 = np.eye(3).reshape(9, 1)
    out = []
    for D in (D1, D2, D1 + D2):
        s.upload(C.FN1, I9 + D)
        out.append(s.f2disp(C.FN1, L))
    lin = np.abs(out[2] - out[0] - out[1]).max()
    print(f"[f2disp] N={N} linearity defect {lin:.3e} (max|u| {np.abs(out[2]).max():.3e})",
          flush=True)
    assert lin <= 1e-10 * max(L)
    # identity: no displacement
    s.upload(C.FN1, np.tile(I9, (1, N ** 3)))
    assert np.abs(s.f2disp(C.FN1, L)).max() <= 1e-13 * max(L)
    s.close()


def test_f2disp_rejects_bad_lengths():
    s = _solver(4)
    with pytest.raises(C.CpfftError):
        s.f2disp(C.FN1, (1.0, 0.0, 1.0))
    s.close()


def test_f2disp_after_load_steps_against_oracle():
    """The deformation gradient the solver leaves in FN1 after real load
    steps (16^3, 8 grains, mixed-BC tension), reconstructed on the GPU and
    by the oracle from the same downloaded field."""
    from cpfft_amd.configs import WORKLOADS
    wl = WORKLOADS["c1_16"]
    BC, isNBC = wl.bc()
    s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    rc, Pbar, _, _ = s.compute(BC[:20], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    assert rc == 0
    F = s.download(C.FN1)
    L = (1.0, 1.0, 1.0)
    u = s.f2disp(C.FN1, L)
    ref = O.f2disp(F, wl.N, L)
    err = np.abs(u - ref).max()
    print(f"\n[f2disp] c1_16 after 20 steps: max|u|={np.abs(ref).max():.3e} max err={err:.3e}",
          flush=True)
    assert np.abs(ref).max() > 1e-3  # a real deformation
    assert err <= 1e-10 * max(L)
    s.close()
