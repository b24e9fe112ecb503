"""Why the workloads load in steps of 2.5e-4 strain (0.25 s at 1e-3/s)
rather than the 1e-3 strain per step the benchmark configs state: the
reference algorithm, run unchanged, does not converge with coarser steps on
these microstructures.  The GPU path here matches the oracle wherever both
run (test_gpu_parity.py), so these runs show the method's behaviour:

* 1 s steps (1e-3 strain): every local solve converges, but the global CG
  of load step 1 does not converge within fftPcg's 1000 iterations once
  the consistent tangent dP/dF of the first plastic update enters the
  operator (the oracle gives the same -2 on this 16^3 case, after six
  minutes of CPU); with 0.25 s steps the same step converges.  The
  smallest eigenvalues of the voxels' symmetric tangents are printed: CG
  needs a positive definite operator on the compatible fields.
* 0.5 s steps at 64^3: the Newton correction flips sign from one iteration
  to the next (voxels at the onset of yield alternate between plastic
  loading and elastic unloading), so |dF|/|F| stays at 2.3e-6 above tolNR
  and the load step ends at the Newton iteration limit (FFT_nr3.f:116-119).
The driver's residual lines (format 1001 of fft_nr3.f90) are captured from
the Fortran unit, and the last two corrections are compared.
"""
import re

import numpy as np
import pytest

import cpfft_amd as C
from cpfft_amd.configs import WORKLOADS, Workload

pytestmark = pytest.mark.gpu


def _tangent_spectrum(s):
    """smallest eigenvalue of the symmetric part of each voxel's dP/dF (K4,
    9x9), as (fraction negative, 1% quantile, minimum) in MPa"""
    K = s.download(C.K4).T.reshape(-1, 9, 9)
    w = np.linalg.eigvalsh(0.5 * (K + K.transpose(0, 2, 1)))[:, 0]
    return float((w < 0).mean()), float(np.quantile(w, 0.01)), float(w.min())


def test_one_second_steps_break_the_global_cg():
    res = {}
    for dt in (1.0, 0.25):
        wl = Workload("dt", 16, 8, 20261202 + 1, 1, 1e-3 * dt, dt=dt)
        BC, isNBC = wl.bc()
        s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
        rc, _, cnt, _ = s.compute(BC, isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
        k = s.counters()
        res[dt] = (rc, k[5], k[6], _tangent_spectrum(s))
        print(f"\n[dt={dt}] rc {rc}: sub-step cuts {k[5]}, failed voxel updates {k[6]}, "
              f"smallest eigenvalue of sym(dP/dF): negative in {res[dt][3][0]:.3f} of the "
              f"voxels, 1% quantile {res[dt][3][1]:.1f} MPa, minimum {res[dt][3][2]:.1f} MPa",
              flush=True)
        s.close()
    assert res[1.0][0] == -2 and res[0.25][0] == 0
    assert res[1.0][2] == 0  # no local solve fails: the global solve does
    # the large increment's tangents are far more indefinite than the
    # stress-sized negative eigenvalues of the small step
    assert res[1.0][3][2] < 20 * res[0.25][3][2] < 0


def test_half_second_steps_cycle_at_64(capfd):
    wl0 = WORKLOADS["c2_64"]
    wl = Workload("dt05", wl0.N, wl0.ngrains, wl0.seed, 4, 2e-3, dt=0.5)
    BC, isNBC = wl.bc()
    last = {}
    # the same run stopped after Newton iteration 24 and after 25: the last
    # corrections dF of the two runs are consecutive iterates of one run
    for m in (24, 25):
        s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
        rc, _, _, _ = s.compute(BC, isNBC, wl.tolPCG, wl.tolNR, m, verbose=True)
        last[m] = (rc, s.download(C.DFM))
        s.close()
    out = capfd.readouterr().out
    res = [float(v) for v in re.findall(r"Newton +\d+ +\|dF\|/\|F\| = +([0-9.Ee+-]+)", out)]
    d24, d25 = last[24][1], last[25][1]
    flip = abs(d25 + d24).max() / abs(d24).max()
    print(f"\n[dt=0.5] rc {last[24][0]}, {last[25][0]}; last Newton residuals {res[-4:]}; "
          f"|dF_25 + dF_24| / |dF_24| = {flip:.2e}", flush=True)
    assert last[24][0] == last[25][0] == -4
    # the correction flips sign from one Newton iteration to the next ...
    assert flip < 1e-2
    # ... so its norm stays fixed above tolNR instead of falling
    tail = np.array(res[-6:])
    assert (tail > wl.tolNR).all() and np.allclose(tail, tail[0], rtol=1e-2)
