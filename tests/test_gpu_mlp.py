"""GPU tests of the warm start of the local solves (no counterpart in the
reference).

* The BF16-MFMA training kernels (forward pass, loss gradient) through the
  C-ABI against the CPU oracle (oracle/mlp_oracle.py), which rounds operands
  to BF16 at the same points, for the fcc (42 -> 12) and bcc24 (66 -> 24)
  networks: 0.2% (output) and 1% per parameter block (gradient), the residue
  being FP32-vs-FP64 accumulation and rare BF16 rounding ties.
* The fused inference kernel against the oracle network on the same
  normalised inputs, for both sizes.
* End to end at 16^3: pairs logged over the first load steps, trained on
  the GPU, then the remaining steps warm-started from the network, from the
  step's previous global iterate, and from both: each reaches the plain
  method's averaged stress (1e-6 relative) with the same global Newton
  iteration counts, none costs more local iterations than the plain start,
  and the previous iterate saves at least a fifth of them.
* Online retraining on a ring log, fcc and bcc24.
"""
import os
import sys

import numpy as np
import pytest

import cpfft_amd as C

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _solver(N=8, ngr=4, seed=20261202 + 99, slip_set=1):
    wl = C.Workload("t", N, ngr, seed, 8, 8 * 2.5e-4, material=C.Material(slip_set=slip_set))
    return C.Solver(N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=1), wl


@pytest.mark.parametrize("n,slip_set", [(500, 1), (2048, 1), (700, 80)])
def test_mlp_forward_and_gradient_match_oracle(n, slip_set):
    s, _ = _solver(slip_set=slip_set)
    s.mlp_init(seed=3, capacity=1024, batch=2048)
    nin, nout, npar, _ = C.mlp_sizes(s.nslip)
    p = s.mlp_params()[:npar]
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, nin)).astype(np.float32)
    T = rng.standard_normal((n, nout)).astype(np.float32)
    Y = s.mlp_forward(X)
    Yo, _ = M.forward(p, X, nout=nout)
    assert np.linalg.norm(Y - Yo) <= 2e-3 * np.linalg.norm(Yo)
    g, loss = s.mlp_grad(X, T)
    lo, go = M.loss_and_grad(p, X, T)
    assert abs(loss - lo) <= 1e-3 * lo
    off = M.offsets(nin, nout)
    for a, b in zip(off[:-1], off[1:]):
        assert np.linalg.norm(g[a:b] - go[a:b]) <= 1e-2 * np.linalg.norm(go[a:b]), (a, b)
    s.close()


def _local(c1, c0):
    """local iterations (predictor + full Newton) between two counter reads"""
    return (c1[3] - c0[3]) + (c1[4] - c0[4])


def test_warm_start_end_to_end():
    """Pairs logged over the first load steps, the network trained on the
    GPU, then the remaining steps run from each warm-start source: the
    network, the step's previous global iterate, and both.  Every source
    reaches the plain method's averaged stress (1e-6 relative) with the same
    global Newton iteration counts; the previous iterate cuts the local
    iterations, and no source costs more than the plain start."""
    N, nlog, nrun = 16, 4, 4
    wl = C.Workload("ws", N, 8, 20261202 + 11, nlog + nrun, (nlog + nrun) * 2.5e-4)
    BC, isNBC = wl.bc()
    ang = wl.angles()
    args = (isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    runs = {}
    for mode in (0, C.WARM_NETWORK, C.WARM_PREVIOUS, C.WARM_NETWORK | C.WARM_PREVIOUS):
        s = C.Solver(N, ang, wl.material, dt=wl.dt, gamma_mode=1)
        if mode & C.WARM_NETWORK:
            s.mlp_init(seed=1, capacity=1 << 16, batch=1024)
            rc, _, _, _ = s.compute(BC[:nlog], *args, mlp_log_last=nlog)
            assert rc == 0 and s.mlp_stats()[0] == nlog * N ** 3
            l0 = s.mlp_train(1)
            loss = s.mlp_train(3000, lr=2e-3)
            assert loss < 0.5 * l0
        else:
            rc, _, _, _ = s.compute(BC[:nlog], *args)
            assert rc == 0
        s.set_warm_start(mode)
        c0 = s.counters()
        rc, P, cnt, times = s.compute(BC[nlog:], *args, first_step=nlog + 1)
        assert rc == 0
        c1 = s.counters()
        runs[mode] = (P, cnt, _local(c1, c0), times[1], s.mlp_stats() if mode & 1 else None)
        s.close()
    P0, cnt0, loc0, t0, _ = runs[0]
    for mode, (P, cnt, loc, t, st) in runs.items():
        rel = abs(P - P0).max() / abs(P0).max()
        print(f"\n[warm {mode}] local iterations {loc} ({loc / loc0:.3f} of plain), "
              f"sig-eps {t:.3f} s, Pbar rel diff {rel:.1e}, newton {cnt[:, 0].tolist()}"
              + (f", redone from plain {st[3]}" if st else ""), flush=True)
        assert rel < 1e-6
        assert np.array_equal(cnt[:, 0], cnt0[:, 0])
        assert loc <= loc0 * 1.02
    assert runs[C.WARM_PREVIOUS][2] < 0.8 * loc0


@pytest.mark.parametrize("slip_set", [1, 80])
def test_online_retraining_on_a_ring_log(slip_set):
    """The surrogate retrained every R load steps on the pairs the method
    logged in the latest steps (the log is a ring smaller than one round of
    pairs, so it wraps), warm-starting the steps in between, for fcc (12
    outputs) and bcc24 (24 outputs).  The FP64 local Newton decides:
    P_bar agrees with the plain run to 1e-6 and the global Newton counts are
    equal."""
    N, R, rounds = 16, 2, 3
    nstep = R * rounds
    wl = C.Workload("online", N, 8, 20261202 + 12, nstep, nstep * 2.5e-4,
                    material=C.Material(slip_set=slip_set))
    BC, isNBC = wl.bc()
    ang = wl.angles()
    args = (isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    a = C.Solver(N, ang, wl.material, dt=wl.dt, gamma_mode=1)
    rc, Pa, cnta, _ = a.compute(BC, *args)
    assert rc == 0
    b = C.Solver(N, ang, wl.material, dt=wl.dt, gamma_mode=1)
    cap = 3 * N ** 3 // 2
    b.mlp_init(seed=1, capacity=cap, batch=512)
    Ps, cnts = [], []
    for k in range(rounds):
        rc, P, cnt, _ = b.compute(BC[k * R:(k + 1) * R], *args, first_step=k * R + 1,
                                  mlp_log_last=(k + 1) * R)
        assert rc == 0
        Ps.append(P)
        cnts.append(cnt)
        st = b.mlp_stats()
        assert st[0] == min((k + 1) * R * N ** 3, cap)
        loss = b.mlp_train(300, lr=2e-3)
        print(f"\n[online] round {k}: pairs held {st[0]}, loss {loss:.4f}", flush=True)
        assert np.isfinite(loss)
        b.mlp_warm_start(True)
    Pb, cntb = np.vstack(Ps), np.vstack(cnts)
    assert abs(Pb - Pa).max() / abs(Pa).max() < 1e-6
    assert np.array_equal(cntb[:, 0], cnta[:, 0])
    assert b.mlp_stats()[2] > 0
    a.close()
    b.close()


@pytest.mark.parametrize("slip_set", [1, 80])
def test_fused_inference_matches_oracle_network(slip_set):
    """The inference path of the warm start (every voxel's inputs, then one
    fused MFMA kernel) against the oracle network on the same normalised
    inputs, for fcc (42 -> 12) and bcc24 (66 -> 24), with trained-shape
    weights and a non-trivial normalisation."""
    N = 8
    wl = C.Workload("inf", N, 4, 20261202 + 31, 2, 5e-4, material=C.Material(slip_set=slip_set))
    s = C.Solver(N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=1)
    BC, isNBC = wl.bc()
    rc, _, _, _ = s.compute(BC[:1], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    assert rc == 0
    s.mlp_init(seed=5, capacity=4096, batch=512)
    nin, nout, npar, nnorm = C.mlp_sizes(s.nslip)
    rng = np.random.default_rng(slip_set)
    p = s.mlp_params()
    p[:npar] += 0.05 * rng.standard_normal(npar).astype(np.float32)  # non-zero biases too
    # the next step's state: an increment on F_n+1
    s.fill9(C.DFM, BC[1] - BC[0])
    s.daxpy(1.0, C.DFM, C.FN1)
    X, _ = s.mlp_predict(2)
    mean_x, sd_x = X.mean(0), X.std(0) + 0.5
    mean_t = rng.standard_normal(nout).astype(np.float32)
    sd_t = (0.5 + rng.random(nout)).astype(np.float32)
    p[npar:] = np.concatenate([mean_x, sd_x, mean_t, sd_t]).astype(np.float32)
    s.mlp_params(p)
    X2, ws = s.mlp_predict(2)
    assert np.array_equal(X, X2)
    Xn = ((X - mean_x.astype(np.float32)) / sd_x.astype(np.float32)).astype(np.float32)
    Yo, _ = M.forward(p[:npar], Xn, nout=nout)
    want = Yo * sd_t + mean_t
    err = np.linalg.norm(ws.T - want) / np.linalg.norm(want)
    print(f"\n[infer] slip_set {slip_set}: {nin} -> {nout}, rel err {err:.2e}", flush=True)
    assert ws.shape == (nout, N ** 3)
    assert err <= 2e-3
    s.close()
