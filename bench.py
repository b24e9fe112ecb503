"""Benchmark: converged load steps per second of the finite-strain FFT
crystal-plasticity solver on one MI355X.

A step is one converged load step of the FFT_nr3 loop (mixed-BC uniaxial
tension of the 128^3, 256-grain FCC polycrystal), driven by the Fortran
driver over the HIP operators.  W warm-up steps run first (they include the
initial elastic call and homogenised tangent), then K steps are timed.
The headline value is the plain method (the reference's local start).
Further solvers repeat the run with the warm start of the local solves on
for the K timed steps: from the step's previous global iterate, and (fcc and
bcc24) from that and the network, trained on the GPU on pairs logged during
the W warm-up steps; their rates, local iteration counts and averaged-stress
differences from the plain run are reported beside the headline.

With --workload c4_batch the run is the batched uneven-RVE workload
instead (24 RVEs of edge 32-96 solved at once; see batch_bench).

Prints one JSON line.  The roofline object describes one CG iteration
(fftPcg): the direction update, the G_K_dF with K4 of the direction and the
x, r update, launched as one group of six kernels (pass A, alpha, passes B,
C, B', pass D); its device time is measured with HIP events on the solver's
stream inside the timed region.  The cpu_baseline object times the oracle (the CPU restatement, all
host cores) on a bounded sample of the same workload and scales it by this
run's operator counts per load step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import cpfft_amd as C  # noqa: E402
from cpfft_amd.configs import BATCH_WORKLOADS, WORKLOADS  # noqa: E402

HBM_PEAK_GBS = 8000.0
# timed steps record HIP events on every TIMING_STRIDE-th launch group of a
# family (coprime to the CG chunk of eight iterations)
TIMING_STRIDE = 7
# The dominant kernel group is one CG iteration on the compile-time-planned
# path with the textbook (symmetric projection) operator, gamma_mode 1, which
# every workload here uses: pass A forms the
# direction p = r + beta p, K4:p and p.(K4:p) = p.q; a single-block kernel
# forms alpha; the three middle passes apply the projection; pass D applies
# x += alpha p and r -= alpha q without storing q and writes the r.r partial
# sums (only the single-block convergence test is outside the group).
# Algorithmic bytes per voxel, counted per CG iteration of fftPcg as G_K_dF
# with K4 (read the direction 72 + read K4 648 + write q 72 = 792) plus five
# 72-byte vector streams of the r, p, x, q updates and dot products.  (The
# fused group itself needs only 1080: read K4, p, r, x, write p, r, x.)
CG_ITERATION_BYTES_PER_VOXEL = 792.0 + 5 * 72.0


def cg_iteration_traffic_per_voxel(N):
    """HBM bytes per voxel the five passes move for that group: pass A reads
    r, p and K4 and writes p and the half spectrum; eight passes over the
    nine-component half spectrum; pass D reads r, x and p and writes r and
    x.  The same count gives 1377 B/voxel for G_K_dF with K4 alone at 128^3,
    which FETCH_SIZE/WRITE_SIZE counters (calibrated on a copy kernel)
    reproduce to 0.1%."""
    nz = N // 2 + 1
    return 72.0 * 2 + 648.0 + 72.0 + 72.0 * 3 + 72.0 * 2 + 8.0 * 144.0 * nz / N


def device_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def cpu_baseline(wl, solver, per_step, budget_s=20.0):
    """Time the oracle on a bounded sample and scale it to one load step:
    (G_K_dF applications per step) x t(G_K_dF) + (drive_eps_sig calls per
    step) x t(drive_eps_sig over all voxels, from a voxel sample)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    V = wl.N ** 3
    # operator sample: the oracle's fftPcg loop at full size on the current
    # state, run for a fixed number of iterations (its tolerance is never
    # met), so the per-iteration time includes the vector updates as in a
    # whole run; one iteration is one G_K_dF with K4
    K4 = solver.download(C.K4)
    fo = O.FFT(wl.N, wl.gamma_mode)
    b = -fo.gkdf(solver.download(C.DFM), K4)
    nit = max(3, min(40, int(10.0 / (1.6e-7 * V))))  # ~10 s at 16 host threads
    t = time.perf_counter()
    _, _, it, _ = fo.pcg(K4, b, 1e-15, maxit=nit)
    t_gkdf = (time.perf_counter() - t) / max(1, it)
    del K4, b
    # material sample: a contiguous slab of voxels at the current state,
    # plastic update (Newton iteration 1)
    nv = min(V, 16384)
    sl = slice(V // 2, V // 2 + nv)
    Fn = np.ascontiguousarray(solver.download(C.FN)[:, sl])
    Fn1 = np.ascontiguousarray(solver.download(C.FN1)[:, sl])
    sn = np.ascontiguousarray(solver.download(C.STATE_N)[:, sl])
    ang = np.ascontiguousarray(wl.angles()[:, sl])
    oc = O.crystal(slip_set=wl.material.slip_set, jac_exact=wl.material.jac_exact)
    O.drive_eps_sig(oc, ang[:, :256], 3, 1, wl.dt, Fn[:, :256], Fn1[:, :256], sn[:, :256])
    t = time.perf_counter()
    O.drive_eps_sig(oc, ang, 3, 1, wl.dt, Fn, Fn1, sn)
    t_des = (time.perf_counter() - t) * (V / nv)
    n_gkdf, n_des = per_step
    step_s = n_gkdf * t_gkdf + n_des * t_des
    return {
        "value": 1.0 / step_s,
        "unit": "load steps/s",
        "cores": threads,
        "kind": "port",
        "implementation": ("oracle/ C restatement with OpenMP: the reference's complex-to-complex "
                           "transform with phase factors, radix-2 for power-of-two N (not MKL), "
                           "the mm10 update voxel by voxel"),
        "sample": (f"oracle fftPcg at {wl.N}^3 for {it} iterations ({t_gkdf:.3f} s per "
                   f"iteration) and oracle "
                   f"drive_eps_sig on {nv} voxels of the current {wl.N}^3 state "
                   f"({t_des * nv / V:.2f} s, x{V / nv:.0f} for all voxels); one load step = "
                   f"{n_gkdf:.0f} G_K_dF + {n_des:.1f} drive_eps_sig calls (this run's counts)"),
        "seconds_per_step": step_s,
    }


def validate_cpu_baseline(name, nstep):
    """The extrapolated CPU baseline against a whole oracle run: the oracle
    runs load steps 1..nstep of the workload (the initial tangent included)
    and is timed end to end, and cpu_baseline's estimate for the same work
    is formed from the same operator counts (the GPU run of those steps,
    identical algorithm)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    wl = WORKLOADS[name]
    BC, isNBC = wl.bc()
    ang = wl.angles()
    s = C.Solver(wl.N, ang, wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    rc, _, cnt, _ = s.compute(BC[:nstep], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    if rc != 0:
        raise SystemExit(f"GPU run failed with code {rc}")
    k = s.counters()  # since creation: G_K_dF applications, drive_eps_sig calls
    est = cpu_baseline(wl, s, (k[0], k[2]))
    s.close()
    print(f"[validate] estimate {est['seconds_per_step']:.1f} s for steps 1-{nstep}; "
          "running the oracle", flush=True)
    oc = O.crystal(slip_set=wl.material.slip_set, jac_exact=wl.material.jac_exact)
    os.environ["ORC_PROGRESS"] = "1"  # one line per Newton iteration
    t = time.perf_counter()
    Po, co, rco, _ = O.fft_nr3(oc, wl.N, ang, wl.dt, BC[:nstep], isNBC, wl.tolPCG, wl.tolNR,
                               wl.maxIter, gamma_mode=wl.gamma_mode)
    measured = time.perf_counter() - t
    return {"workload": f"{name}, load steps 1-{nstep} (initial tangent included)",
            "cores": est["cores"], "oracle_rc": rco,
            "measured_s": measured, "extrapolated_s": est["seconds_per_step"],
            "measured_over_extrapolated": measured / est["seconds_per_step"],
            "G_K_dF_applications": k[0], "drive_eps_sig_calls": k[2],
            "gpu_counters_per_step": cnt.tolist(), "oracle_counters_per_step": co.tolist(),
            "sample": est["sample"]}


def workload_text(wl):
    """Crystal and loading of a workload, e.g. 'FCC cubic Voce n=20, uniaxial
    tension 1e-3/s'."""
    crystal = {1: "FCC", 2: "BCC", 6: "FCC (roters order)", 7: "BCC12", 8: "BCC48",
               80: "BCC24"}[wl.material.slip_set]
    span = wl.strain * (4 * wl.cycles if wl.cycles > 0 else 1)  # strain travelled
    rate = f"{span / (wl.nstep * wl.dt):.0e}".replace("e-0", "e-")
    load = (f"uniaxial tension-compression +-{100 * wl.strain:g}% x{wl.cycles}, {rate}/s"
            if wl.cycles > 0 else f"uniaxial tension {rate}/s")
    return f"{crystal} cubic Voce n=20, {load}"


def timed_steps(s, wl, BC_all, isNBC, W, K):
    """Time load steps W+1..W+K on a solver that has run steps 1..W;
    returns (wall, rc, Pbar, cnt, times, counter deltas, kernel avg ms, counts)."""
    c0 = s.counters()
    # HIP events on every 7th launch group of each family: each recorded
    # pair idles the stream for about 10 us, so timing every group would
    # slow the steps being timed
    s.set_timing(TIMING_STRIDE)
    device_sync()
    s.sync()
    t0 = time.perf_counter()
    rc, Pbar, cnt, times = s.compute(BC_all[W:W + K], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter,
                                     first_step=W + 1)
    s.sync()
    device_sync()
    t1 = time.perf_counter()
    avg, nlaunch = s.kernel_times()
    c1 = s.counters()
    return t1 - t0, rc, Pbar, cnt, times, [b - a for a, b in zip(c0, c1)], avg, nlaunch


def warm_start_leg(wl, ang, BC_all, isNBC, W, K, Pbar_plain, d_plain, mode):
    """The same W + K steps with the warm start of the local solves on for
    the K timed steps (cpfft_set_warm_start mode).  With the network source
    the network is trained on pairs logged during the W warm-up steps."""
    V = wl.N ** 3
    s = C.Solver(wl.N, ang, wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    s.autotune()  # the shapes chosen for the plain run (cached per N)
    training = None
    if mode & C.WARM_NETWORK:
        stride = max(1, (W * V) // (1 << 21))  # at most ~2M logged pairs
        s.mlp_init(seed=1, capacity=W * ((V + stride - 1) // stride), batch=4096)
        rc, _, _, _ = s.compute(BC_all[:W], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter, first_step=1,
                                mlp_log_last=W, mlp_log_stride=stride)
        if rc != 0:
            raise SystemExit(f"warm-start warm-up failed with code {rc}")
        t = time.perf_counter()
        iters = 3000
        loss = s.mlp_train(iters, lr=2e-3)
        training = {"pairs": int(s.mlp_stats()[0]), "log_stride": stride, "adam_steps": iters,
                    "seconds": time.perf_counter() - t, "final_mse_normalised": loss}
    else:
        rc, _, _, _ = s.compute(BC_all[:W], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter, first_step=1)
        if rc != 0:
            raise SystemExit(f"warm-start warm-up failed with code {rc}")
    s.set_warm_start(mode)
    l0 = s.local_stats()
    wall, rc, Pbar, cnt, times, d, avg, _ = timed_steps(s, wl, BC_all, isNBC, W, K)
    l1 = s.local_stats()
    if rc != 0:
        raise SystemExit(f"warm-started steps failed with code {rc}")
    st = s.mlp_stats()
    s.close()
    n_des = d[2]
    lw = (d[3] + d[4]) / max(1, n_des * V)
    lp = (d_plain[3] + d_plain[4]) / max(1, d_plain[2] * V)
    out = {
        "mode": {C.WARM_NETWORK: "network", C.WARM_PREVIOUS: "previous iterate",
                 C.WARM_NETWORK | C.WARM_PREVIOUS: "network and previous iterate"}[mode],
        "value": K / wall, "unit": "load steps/s", "ms_per_step": 1e3 * wall / K,
        "voxel_updates_per_s": V * n_des / wall,
        "drive_eps_sig_ms": avg[1],
        "local_iterations_per_voxel_update": lw,
        "plain_local_iterations_per_voxel_update": lp,
        "local_iteration_lane_efficiency": (l1[0] - l0[0]) / max(1, l1[1] - l0[1]),
        "Pbar_max_rel_diff_vs_plain": float(abs(Pbar - Pbar_plain).max() / abs(Pbar_plain).max()),
        "newton_per_step": float(cnt[:, 0].mean()),
        "redone_from_plain_start": int(st[3]),
    }
    if training:
        out["training"] = training
    return out


def online_bench(name, W, R, M):
    """Online retraining at full size: after W plain steps, M rounds of R
    load steps; every round logs training pairs, and before each round after
    the first the network is retrained on the latest logged pairs (a ring of
    one round's pairs) and warm-starts the round's local solves.  A plain run
    of the same steps is the reference.  Reports per round the local
    iterations, the time and the averaged-stress difference."""
    wl = WORKLOADS[name]
    BC, isNBC = wl.bc()
    ang = wl.angles()
    V = wl.N ** 3
    args = (isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
    legs = {}
    for leg in ("plain", "online"):
        s = C.Solver(wl.N, ang, wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
        s.autotune()
        if leg == "online":
            stride = max(1, (R * V) // (1 << 21))  # at most ~2M pairs a round
            s.mlp_init(seed=1, capacity=R * ((V + stride - 1) // stride), batch=4096)
        rc, _, _, _ = s.compute(BC[:W], *args)
        if rc != 0:
            raise SystemExit(f"{leg}: warm-up failed with code {rc}")
        rounds = []
        for k in range(M):
            first = W + k * R + 1
            info = {"steps": [first, first + R - 1]}
            if leg == "online" and k > 0:
                t = time.perf_counter()
                info["mse_normalised"] = s.mlp_train(1000, lr=2e-3)
                info["train_s"] = time.perf_counter() - t
                s.set_warm_start(C.WARM_NETWORK)
            c0 = s.counters()
            t = time.perf_counter()
            kw = {"mlp_log_last": first + R - 1, "mlp_log_stride": stride} if leg == "online" else {}
            rc, P, cnt, _ = s.compute(BC[first - 1:first - 1 + R], *args, first_step=first, **kw)
            info["wall_s"] = time.perf_counter() - t
            if rc != 0:
                raise SystemExit(f"{leg}: round {k} failed with code {rc}")
            c1 = s.counters()
            info["local_iterations_per_voxel_update"] = (
                (c1[3] - c0[3]) + (c1[4] - c0[4])) / max(1, (c1[2] - c0[2]) * V)
            info["newton"] = cnt[:, 0].tolist()
            info["P_xx"] = P[:, 0].tolist()
            rounds.append((info, P))
            print(f"[online] {leg} round {k}: {json.dumps(info)}", flush=True)
        legs[leg] = rounds
        if leg == "online":
            legs["stats"] = s.mlp_stats()
        s.close()
    out = {"workload": f"{name}: {wl.N}^3, {wl.ngrains} grains, {workload_text(wl)}, dt={wl.dt}s",
           "warmup_steps": W, "round_steps": R, "rounds": M, "network_pairs_per_round": None,
           "rounds_plain": [i for i, _ in legs["plain"]],
           "rounds_online": [i for i, _ in legs["online"]]}
    st = legs["stats"]
    out["network_pairs_held"] = int(st[0])
    out["warm_started_voxel_updates"] = int(st[2])
    out["redone_from_plain_start"] = int(st[3])
    out["Pbar_max_rel_diff_vs_plain"] = max(
        float(abs(Po - Pp).max() / abs(Pp).max())
        for (_, Pp), (_, Po) in zip(legs["plain"], legs["online"]))
    del out["network_pairs_per_round"]
    return out


def batch_bench(name, W, K, nthreads):
    """Batched uneven RVEs: the members' load steps W+1..W+K timed twice
    from the same warm-up state, once one RVE after another on one host
    thread with kernel-by-kernel launches and the default FFT launch shapes,
    once through cpfft_compute_batch on nthreads threads with autotuned
    shapes and graph-replayed CG chunks.  Returns the JSON record (value:
    the batched aggregate rate)."""
    bw = BATCH_WORKLOADS[name]
    members = bw.members()
    angles = [w.angles() for w in members]
    bcs = [w.bc() for w in members]
    vox = sum(w.N ** 3 for w in members)
    print(f"[bench] {name}: {len(members)} RVEs, edges {[w.N for w in members]}, "
          f"{vox} voxels in all, warmup={W} steps={K}", flush=True)
    legs = {}
    shapes = {}
    for leg, nt, graphs, tuned in (("sequential", 1, False, False),
                                   ("batched", nthreads, True, True)):
        solvers = []
        for w, ang in zip(members, angles):
            sv = C.Solver(w.N, ang, w.material, dt=w.dt, gamma_mode=w.gamma_mode)
            sv.set_graphs(graphs)
            if tuned:
                shapes[w.N] = sv.autotune()  # measured once per edge, then cached
            solvers.append(sv)

        def jobs(lo, hi, first):
            return [(sv, bc[0][lo:hi], bc[1], w.tolPCG, w.tolNR, w.maxIter, first)
                    for sv, w, bc in zip(solvers, members, bcs)]
        rc, _ = C.compute_batch(jobs(0, W, 1), nthreads=nt)
        if rc != 0:
            raise SystemExit(f"{leg} warm-up failed with code {rc}")
        c0 = [sv.counters() for sv in solvers]
        device_sync()
        t0 = time.perf_counter()
        rc, res = C.compute_batch(jobs(W, W + K, W + 1), nthreads=nt)
        for sv in solvers:
            sv.sync()
        device_sync()
        wall = time.perf_counter() - t0
        if rc != 0:
            raise SystemExit(f"{leg} timed steps failed with code {rc}")
        c1 = [sv.counters() for sv in solvers]
        legs[leg] = {
            "wall_s": wall,
            "voxel_load_steps_per_s": vox * K / wall,
            "rve_load_steps_per_s": len(members) * K / wall,
            "voxel_updates_per_s": sum(w.N ** 3 * (b[2] - a[2])
                                       for w, a, b in zip(members, c0, c1)) / wall,
            "cg_iterations": int(sum(b[1] - a[1] for a, b in zip(c0, c1))),
            "threads": nt, "graphs": graphs, "autotuned": tuned,
            "Pbar": [r[1] for r in res],
            "newton": [int(r[2][:, 0].sum()) for r in res],
        }
        print(f"[bench] {leg}: {wall:.2f} s, {vox * K / wall:.3e} voxel load steps/s", flush=True)
        for sv in solvers:
            sv.close()
    seq, bat = legs["sequential"], legs["batched"]
    same = all(np.array_equal(a, b) for a, b in zip(seq.pop("Pbar"), bat.pop("Pbar")))
    return {
        "metric": f"aggregate voxel load steps/s over {len(members)} uneven RVEs",
        "value": bat["voxel_load_steps_per_s"],
        "unit": "voxel load steps/s",
        "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": 1e3 * bat["wall_s"] / K,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": (f"{name}: {len(members)} periodic Voronoi RVEs, edges {sorted(set(bw.edges))}, "
                         f"{bw.grains[0]}-{bw.grains[1]} grains, FCC cubic Voce n=20, F = I + eps n(x)n "
                         f"along random n, {bw.dstrain} strain per step at rates "
                         f"{bw.rates[0]}-{bw.rates[1]}/s, steps {W + 1}-{W + K}"),
            "edges": [w.N for w in members], "grains": [w.ngrains for w in members],
            "dt": [w.dt for w in members], "voxels": vox,
            "fft_launch_shapes": {str(k): v for k, v in sorted(shapes.items())},
        },
        "batched": bat, "sequential": seq,
        "speedup_vs_sequential": seq["wall_s"] / bat["wall_s"],
        "Pbar_bitwise_equal_to_sequential": bool(same),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3_128")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true")
    ap.add_argument("--threads", type=int, default=8, help="host threads of a batched workload")
    ap.add_argument("--online", type=int, nargs=2, default=None, metavar=("R", "M"),
                    help="online retraining run of --workload: M rounds of R load steps "
                         "after --warmup plain steps, then exit")
    ap.add_argument("--validate-cpu-baseline", type=int, default=0, metavar="NSTEP",
                    help="time a whole oracle run of NSTEP load steps of --workload against "
                         "the extrapolated CPU baseline, then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the solver step is built and measured on one GPU's share)", file=sys.stderr)
        sys.exit(2)
    if args.validate_cpu_baseline > 0:
        print(json.dumps(validate_cpu_baseline(args.workload, args.validate_cpu_baseline)),
              flush=True)
        return
    if args.online:
        print(json.dumps(online_bench(args.workload, args.warmup, *args.online)), flush=True)
        return
    if args.workload in BATCH_WORKLOADS:
        print(json.dumps(batch_bench(args.workload, args.warmup, args.steps, args.threads)),
              flush=True)
        return

    wl = WORKLOADS[args.workload]
    W, K = args.warmup, args.steps
    nstep = W + K
    BC_all, isNBC = wl.bc()
    if nstep > BC_all.shape[0]:
        raise SystemExit(f"{args.workload} has {BC_all.shape[0]} load steps")
    ang = wl.angles()
    s = C.Solver(wl.N, ang, wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    shapes = s.autotune()
    print(f"[bench] {wl.name}: N={wl.N} grains={wl.ngrains} warmup={W} steps={K} "
          f"FFT launch shapes {shapes}", flush=True)

    rc, Pw, cw, _ = s.compute(BC_all[:W], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter, first_step=1)
    if rc != 0:
        raise SystemExit(f"warm-up failed with code {rc}")
    print(f"[bench] warm-up done: counters {cw.tolist()}", flush=True)
    l0 = s.local_stats()
    wall, rc, Pbar, cnt, times, d, avg, nlaunch = timed_steps(s, wl, BC_all, isNBC, W, K)
    l1 = s.local_stats()
    if rc != 0:
        raise SystemExit(f"timed steps failed with code {rc}")
    V = wl.N ** 3
    n_des = d[2]
    n_gkdf = d[0]
    steps_per_s = K / wall
    gk_ms = avg[0]
    achieved = CG_ITERATION_BYTES_PER_VOXEL * V / (gk_ms * 1e-3) / 1e9 if gk_ms > 0 else 0.0
    out = {
        "metric": f"converged load steps/s at {wl.N}^3 (plain method; warm start in warm_start)",
        "value": steps_per_s,
        "unit": "load steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{wl.name}: {wl.N}^3 periodic Voronoi, {wl.ngrains} grains, "
                         f"{workload_text(wl)}, dt={wl.dt}s, steps {W + 1}-{nstep}"),
            "grid": wl.N, "grains": wl.ngrains, "gamma_mode": wl.gamma_mode,
            "jac_exact": wl.material.jac_exact, "tolPCG": wl.tolPCG, "tolNR": wl.tolNR,
            "fft_launch_shapes": shapes,
        },
        "voxel_updates_per_s": V * n_des / wall,
        "per_step": {
            "G_K_dF": n_gkdf / K, "cg_iterations": d[1] / K, "drive_eps_sig": n_des / K,
            "newton": float(cnt[:, 0].mean()), "mixed_bc_loops": float(cnt[:, 2].mean()),
            "local_newton_iters_per_voxel_call": d[4] / max(1, n_des * V),
            "local_failures": d[6],
            # local iterations of the voxels / iterations their wavefronts
            # issue (64 x the slowest voxel of each wavefront)
            "local_iteration_lane_efficiency": (l1[0] - l0[0]) / max(1, l1[1] - l0[1]),
        },
        "category_s": {"pcg": float(times[0]), "sig_eps": float(times[1]), "total": float(times[2])},
        "kernel_ms": {"cg_iteration_group": avg[0], "drive_eps_sig": avg[1], "cg_check": avg[2],
                      "G_K_dF_no_K4": avg[3]},
        "roofline": {
            "kernel": ("CG iteration: p = r + beta p and G_K_dF with K4 and p.q (pass A), "
                       "alpha, projection, x and r update with r.r (pass D); 6-kernel launch group"),
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": cg_iteration_traffic_per_voxel(wl.N) * V,
            "algorithmic_bytes_per_launch": CG_ITERATION_BYTES_PER_VOXEL * V,
            "avg_launch_ms": gk_ms,
            "launches_measured": int(nlaunch[0]),
        },
        "Pbar_xx_last": float(Pbar[-1, 0]),
    }
    # strain covered per second, and the rate in the benchmark config's own
    # unit of 1e-3 strain per load step (its load steps are 4x those here)
    dstrain = float(BC_all[1, 0] - BC_all[0, 0]) if wl.cycles == 0 else wl.strain * 4 * wl.cycles / wl.nstep
    out["strain_per_step"] = dstrain
    out["strain_per_s"] = steps_per_s * dstrain
    out["config_step_equivalent"] = {
        "value": steps_per_s * dstrain / 1e-3, "unit": "load steps of 1e-3 strain per s",
        "note": ("the configs' 1e-3 strain per step does not converge with the reference "
                 "algorithm (tests/test_gpu_step_size.py); this run covers the same strain "
                 "in steps of {:g}".format(dstrain))}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, s, (n_gkdf / K, n_des / K))
    s.close()
    out["per_step"]["drive_eps_sig_ms"] = avg[1]
    if not args.no_warm_start:
        modes = [C.WARM_PREVIOUS]
        if wl.material.slip_set != 8:  # the network serves up to 24 slip systems
            modes.append(C.WARM_NETWORK | C.WARM_PREVIOUS)
        out["warm_start"] = []
        for mode in modes:
            print(f"[bench] warm-start leg, mode {mode}", flush=True)
            out["warm_start"].append(warm_start_leg(wl, ang, BC_all, isNBC, W, K, Pbar, d, mode))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
