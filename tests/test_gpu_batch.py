"""GPU tests of the execution options that have no counterpart in the
reference: CG chunks replayed from HIP graphs (cpfft_set_graphs) and
batched uneven RVEs (cpfft_compute_batch).  Both run the same kernels in
the same order on each RVE's stream as the eager single-RVE path, so the
results must agree bit for bit with it (the eager path itself is checked
against the oracle in test_gpu_parity.py).  Graphs are replayed many times
per solve, with the solver's own eager kernels and host synchronisations in
between, and on new data in every call."""
import numpy as np
import pytest

import cpfft_amd as C
from cpfft_amd.microstructure import polycrystal

pytestmark = pytest.mark.gpu


def _solver(N, seed=5, mode=1, dt=0.25):
    _, ang = polycrystal(N, 4, seed)
    return C.Solver(N, ang, C.Material(), dt=dt, gamma_mode=mode)


@pytest.mark.parametrize("N,mode", [(12, 0), (16, 1)])
def test_graph_replay_matches_eager_bitwise(N, mode):
    """N = 12: runtime-plan FFT path; N = 16: compile-time-planned path
    with the CG work fused into the operator."""
    a, b = _solver(N, mode=mode), _solver(N, mode=mode)
    b.set_graphs(True)
    for s in (a, b):
        s.drive_eps_sig(1, 0)
    rng = np.random.default_rng(N)
    for rep in range(4):
        dF = 1e-3 * rng.standard_normal((9, N ** 3))
        x = C.DFM if rep % 2 else C.WORK  # two solution fields: two graphs
        out = []
        for s in (a, b):
            s.upload(C.FN1, dF)
            s.gkdf(C.FN1, C.B, 1)
            s.dscal(-1.0, C.B)
            rc, it, rel = s.pcg(C.B, x, 1e-8)
            out.append((rc, it, rel, s.download(x)))
        print(f"\n[graph] N={N} rep={rep} CG iterations {out[0][1]}", flush=True)
        assert out[0][0] == out[1][0] == 0 and out[0][1] == out[1][1] > 8
        assert out[0][2] == out[1][2]
        assert np.array_equal(out[0][3], out[1][3])
    Ca = a.tangent_homo(1e-8)
    Cb = b.tangent_homo(1e-8)
    assert a.counters()[1] == b.counters()[1]  # same CG iteration totals
    Cb2 = b.tangent_homo(1e-8)
    assert np.array_equal(Ca, Cb) and np.array_equal(Cb, Cb2)
    a.close()
    b.close()


def test_load_steps_with_graphs_match_eager_bitwise():
    N, nstep = 16, 3
    wl = C.Workload("g", N, 4, 20261202 + 98, nstep, 2.5e-4 * nstep)
    BC, isNBC = wl.bc()
    res = []
    for g in (False, True):
        s = C.Solver(N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
        s.set_graphs(g)
        rc, Pbar, cnt, _ = s.compute(BC, isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
        assert rc == 0
        res.append((Pbar, cnt, s.counters()))
        s.close()
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]


def _uneven_jobs():
    """Four RVEs of different grid edge (runtime-plan and planned FFT
    paths), grain count, time step and loading."""
    spec = [(16, 6, 0.25, 3, 7.5e-4), (12, 4, 0.5, 2, 1e-3), (20, 8, 0.25, 2, 5e-4),
            (16, 3, 0.125, 4, 5e-4)]
    jobs = []
    for k, (N, ngr, dt, nstep, strain) in enumerate(spec):
        wl = C.Workload(f"b{k}", N, ngr, 20261202 + 40 + k, nstep, strain, dt=dt)
        BC, isNBC = wl.bc()
        jobs.append((wl, BC, isNBC))
    return jobs


def test_batch_matches_sequential_bitwise():
    jobs = _uneven_jobs()
    seq = []
    for wl, BC, isNBC in jobs:
        s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
        rc, Pbar, cnt, _ = s.compute(BC, isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
        assert rc == 0
        seq.append((Pbar, cnt))
        s.close()
    # graphs: 0 off, 1 captured before the batch starts, 2 captured lazily at
    # first use while other threads run their RVEs (captures are serialised
    # and each is checked to hold exactly the chunk's kernels)
    for graphs in (0, 1, 2):
        solvers = []
        batch = []
        for wl, BC, isNBC in jobs:
            s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
            s.set_graphs(graphs)
            solvers.append(s)
            batch.append((s, BC, isNBC, wl.tolPCG, wl.tolNR, wl.maxIter, 1))
        rc, res = C.compute_batch(batch, nthreads=4)
        print(f"\n[batch] graphs={graphs} rc={rc} rcs={[r[0] for r in res]}", flush=True)
        assert rc == 0
        for (rck, Pbar, cnt, times), (Ps, cs) in zip(res, seq):
            assert rck == 0
            assert np.array_equal(Pbar, Ps)
            assert np.array_equal(cnt, cs)
            assert times[2] > 0
        for s in solvers:
            s.close()


def test_batch_continues_runs():
    """Driver state lives in each context: two batches of steps on the same
    contexts equal one run of all the steps."""
    jobs = _uneven_jobs()[:2]
    full, split = [], []
    for wl, BC, isNBC in jobs:
        s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
        rc, Pbar, _, _ = s.compute(BC, isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
        assert rc == 0
        full.append(Pbar)
        s.close()
    solvers = [C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
               for wl, _, _ in jobs]
    rc1, r1 = C.compute_batch([(s, BC[:1], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter, 1)
                               for s, (wl, BC, isNBC) in zip(solvers, jobs)], nthreads=2)
    rc2, r2 = C.compute_batch([(s, BC[1:], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter, 2)
                               for s, (wl, BC, isNBC) in zip(solvers, jobs)], nthreads=2)
    assert rc1 == rc2 == 0
    for k in range(len(jobs)):
        assert np.array_equal(np.vstack([r1[k][1], r2[k][1]]), full[k])
    for s in solvers:
        s.close()


@pytest.mark.parametrize("N", [16, 32, 48, 64, 96, 128, 256])
def test_every_tuned_shape_gives_identical_results(N):
    """The autotuner's candidates (cpfft_set_tune accepts exactly those)
    change only which block transforms which lines, and for N = 64, 128 and
    256 whether pass C runs its transforms in LDS or in registers: G_K_dF with
    and without K4 and a fused CG solve are bitwise equal under every one of
    them."""
    s = _solver(N)
    s.drive_eps_sig(1, 0)
    rng = np.random.default_rng(3 * N)
    F = rng.standard_normal((9, N ** 3))
    default = s.tune()
    assert all(v > 0 for v in default)

    def run():
        s.upload(C.DFM, F)
        s.gkdf(C.DFM, C.B, 0)
        g0 = s.download(C.B)
        s.gkdf(C.DFM, C.B, 1)
        g1 = s.download(C.B)
        s.dscal(-1e-3, C.B)
        rc, it, _ = s.pcg(C.B, C.WORK, 1e-8)
        assert rc == 0
        return g0, g1, it, s.download(C.WORK)

    ref = run()
    tried = 0
    for k in range(4):
        for v in range(1, 65):
            shape = list(default)
            shape[k] = v
            if v == default[k]:
                continue
            try:
                s.set_tune(shape)
            except C.CpfftError:
                continue
            assert s.tune() == shape
            print(f"\n[tune] N={N} shape {shape}", flush=True)
            got = run()
            tried += 1
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
            assert got[2] == ref[2] and np.array_equal(got[3], ref[3])
    assert tried >= 3
    chosen = s.autotune(reps=3, force=True)
    print(f"\n[tune] N={N} default {default} autotuned {chosen}", flush=True)
    got = run()
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[3], ref[3])
    s.close()


def test_sampled_kernel_timing():
    """cpfft_set_timing(k) records every k-th launch group; the iterates do
    not depend on it and the no-op iterations after convergence are not
    counted."""
    N = 16
    s = _solver(N)
    s.drive_eps_sig(1, 0)
    rng = np.random.default_rng(11)
    s.upload(C.DFM, 1e-3 * rng.standard_normal((9, N ** 3)))
    out = []
    for k in (0, 1, 3):
        s.gkdf(C.DFM, C.B, 1)
        s.dscal(-1.0, C.B)
        s.set_timing(k)
        rc, it, _ = s.pcg(C.B, C.WORK, 1e-8)
        avg, cnt = s.kernel_times()
        print(f"\n[timing] stride {k}: {it} iterations, groups timed {list(cnt)}", flush=True)
        assert rc == 0
        want = 0 if k == 0 else -(-it // k)
        assert cnt[0] == want and cnt[2] == want
        assert (avg[0] > 0) == (k > 0)
        out.append((it, s.download(C.WORK)))
    assert out[0][0] == out[1][0] == out[2][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][1], out[2][1])
    with pytest.raises(C.CpfftError):
        s.set_timing(-1)
    s.close()
