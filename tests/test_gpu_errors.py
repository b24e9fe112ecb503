"""GPU tests of the solver's failure paths, each forced from the host so
every kernel still exits normally:

* fftPcg's iteration limit (FFT_nr3.f:224, 335-339): a solve that needs
  exactly maxIter iterations fails even though its last iteration passed
  the convergence test (itercount >= maxIter aborts); one more allowed
  iteration succeeds with the same iterate.
* the same limit in tangent_homo's nine solves;
* an improper CG tolerance stops the load-step driver before any step
  (fftPcg's die_abort, FFT_nr3.f:248-251), also in a continued run whose
  steps would otherwise never reach a CG solve with a nonzero right-hand
  side.
* the warm-start mode's argument checks, and the local-solve counters.
"""
import numpy as np
import pytest

import cpfft_amd as C
from cpfft_amd.microstructure import polycrystal

pytestmark = pytest.mark.gpu


def _solver(N=8, mode=1):
    _, ang = polycrystal(N, 4, 17)
    s = C.Solver(N, ang, C.Material(), dt=0.25, gamma_mode=mode)
    s.drive_eps_sig(1, 0)
    return s


@pytest.mark.parametrize("N", [8, 16])
def test_cg_fails_at_the_iteration_limit(N):
    s = _solver(N)
    rng = np.random.default_rng(N)
    s.upload(C.DFM, 1e-3 * rng.standard_normal((9, N ** 3)))
    s.gkdf(C.DFM, C.B, 1)
    s.dscal(-1.0, C.B)
    rc, k, _ = s.pcg(C.B, C.WORK, 1e-8)
    assert rc == 0 and k > 2
    x = s.download(C.WORK)
    print(f"\n[limit] N={N}: {k} iterations", flush=True)
    s.set_cg_maxit(k)
    with pytest.raises(C.CpfftError):
        s.pcg(C.B, C.WORK, 1e-8)
    assert np.array_equal(s.download(C.WORK), x)  # the same iterate, reported as failed
    s.set_cg_maxit(k + 1)
    rc, k2, _ = s.pcg(C.B, C.WORK, 1e-8)
    assert rc == 0 and k2 == k and np.array_equal(s.download(C.WORK), x)
    s.set_cg_maxit(1000)
    s.tangent_homo(1e-8)
    s.set_cg_maxit(2)
    with pytest.raises(C.CpfftError):
        s.tangent_homo(1e-8)
    with pytest.raises(C.CpfftError):
        s.set_cg_maxit(0)
    s.close()


def test_driver_stops_on_an_improper_cg_tolerance():
    N = 8
    wl = C.Workload("tol", N, 4, 20261202 + 77, 2, 5e-4)
    BC, isNBC = wl.bc()
    s = C.Solver(N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    for tol in (0.0, 1.0, 1e-17):
        rc, _, cnt, _ = s.compute(BC, isNBC, tol, wl.tolNR, wl.maxIter)
        assert rc == -1 and not cnt.any()
    rc, P1, _, _ = s.compute(BC[:1], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    assert rc == 0
    # a continued run with every component prescribed and no increment: the
    # right-hand sides are zero, yet the bad tolerance is still reported
    allF = np.zeros(9, np.int32)
    Fb = np.tile(s.driver_state()[0], (1, 1))
    rc, _, _, _ = s.compute(Fb, allF, 2.0, wl.tolNR, wl.maxIter, first_step=2)
    assert rc == -1
    s.close()


def test_warm_start_mode_errors_and_local_stats():
    """cpfft_set_warm_start rejects unknown bits and a network start without
    a network; the previous-iterate start needs no network.  The local-solve
    counters: the iterations wavefronts issue are at least the voxels' own."""
    s = _solver(8)
    with pytest.raises(RuntimeError):
        s.set_warm_start(4)
    with pytest.raises(RuntimeError):
        s.set_warm_start(C.WARM_NETWORK)
    s.set_warm_start(C.WARM_PREVIOUS)
    s.set_warm_start(0)
    l0 = s.local_stats()
    s.fill9(C.DFM, np.array([5e-4, 0, 0, 0, -2.5e-4, 0, 0, 0, -2.5e-4]))
    s.daxpy(1.0, C.DFM, C.FN1)
    assert s.drive_eps_sig(1, 1) == 0
    l1 = s.local_stats()
    useful, issued = l1[0] - l0[0], l1[1] - l0[1]
    print(f"\n[local] iterations {useful}, issued {issued}", flush=True)
    assert useful > 0 and issued >= useful
    assert issued % 64 == 0
    s.close()
