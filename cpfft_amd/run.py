"""Run a named workload (configs.WORKLOADS) through its load steps on the GPU,
the way the reference's deck command `compute` runs a loading history
(FFT_nr3.f:51-185 through the Fortran driver): steps `--first` to `--last`,
`--group` steps per cpfft_compute call, a progress line after each group
and a summary line at the end.

    python -m cpfft_amd.run --workload c3_128                # all 240 steps
    python -m cpfft_amd.run --workload c5_256 --last 64 --group 1
    python -m cpfft_amd.run --workload c3_128 --warm-previous

A run stops at the first step the driver fails (rc < 0), as the reference
stops at its die_abort paths.
"""
import argparse
import json
import time

import numpy as np

from . import WARM_PREVIOUS, Solver
from .configs import WORKLOADS


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="c3_128", choices=sorted(WORKLOADS))
    ap.add_argument("--first", type=int, default=1, help="first load step (1 = from the start)")
    ap.add_argument("--last", type=int, default=0, help="last load step (0 = the workload's last)")
    ap.add_argument("--group", type=int, default=10, help="load steps per cpfft_compute call")
    ap.add_argument("--no-graphs", action="store_true", help="launch the CG chunks eagerly")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--warm-previous", action="store_true",
                    help="start the local solves from the step's previous global iterate "
                         "(cpfft_set_warm_start with CPFFT_WARM_PREVIOUS)")
    args = ap.parse_args(argv)
    wl = WORKLOADS[args.workload]
    last = args.last or wl.nstep
    if not 1 <= args.first <= last <= wl.nstep or args.group < 1:
        ap.error(f"steps must satisfy 1 <= first <= last <= {wl.nstep}, group >= 1")
    BC, isNBC = wl.bc()
    t0 = time.perf_counter()
    s = Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    if not args.no_graphs:
        s.set_graphs(True)
    if not args.no_autotune:
        s.autotune()
    if args.warm_previous:
        s.set_warm_start(WARM_PREVIOUS)
    print(f"[run] {wl.name}: N={wl.N}, {wl.ngrains} grains, steps {args.first}-{last}; "
          f"set up in {time.perf_counter() - t0:.1f} s, shapes {s.tune()}", flush=True)
    if args.first > 1:
        # the history before `first`, not timed
        rc, _, _, _ = s.compute(BC[:args.first - 1], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter)
        if rc != 0:
            raise SystemExit(f"[run] the steps before {args.first} stopped with rc {rc}")
    loc0 = s.local_stats()
    t1 = time.perf_counter()
    counts, rc, done = [], 0, args.first - 1
    for k0 in range(args.first - 1, last, args.group):
        k1 = min(k0 + args.group, last)
        rc, P, cnt, _ = s.compute(BC[k0:k1], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter,
                                  first_step=k0 + 1)
        ok = cnt.any(axis=1)  # a step's counts are written when it ends
        counts.append(cnt[ok])
        done = k0 + int(np.count_nonzero(ok))
        eps = BC[done - 1, 0] - 1.0 if done > 0 else 0.0
        print(f"[run] steps {k0 + 1}-{k1}: rc {rc}, F_xx - 1 = {eps:+.5f}, "
              f"P_xx {P[ok][-1, 0] if ok.any() else float('nan'):+.4f}, CG per step "
              f"{cnt[ok, 1].mean() if ok.any() else 0:.0f}, {time.perf_counter() - t1:.1f} s",
              flush=True)
        if rc != 0:
            break
    wall = time.perf_counter() - t1
    c = np.vstack(counts) if counts else np.zeros((0, 4), np.int32)
    n = c.shape[0]
    loc = s.local_stats()
    print(json.dumps({
        "workload": wl.name, "steps": [args.first, done], "rc": int(rc), "seconds": wall,
        "warm_start": "previous iterate" if args.warm_previous else None,
        "local_iterations": int(loc[0] - loc0[0]),  # of the voxels, over the timed steps
        "load_steps_per_s": n / wall if wall > 0 else None,
        "per_step": {k: float(c[:, j].mean()) if n else None for j, k in
                     enumerate(("newton", "cg_iterations", "mixed_bc_loops", "drive_eps_sig"))},
    }), flush=True)
    s.close()
    return 0 if rc == 0 else 1


if __name__ == "__main__":
    raise SystemExit(main())
