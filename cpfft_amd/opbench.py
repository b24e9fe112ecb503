"""Operator micro-benchmark: device time of each hot operator of the solver
step on one RVE, from HIP events recorded on the solver's stream.

    python -m cpfft_amd.opbench [--workload c3_128] [--reps 40]

Operators (the reference routines they replace):
  G_K_dF with K4     G_K_dF.f:11-19, flgK = 1 (the CG matrix-vector product)
  G_K_dF without K4  G_K_dF.f:11-19, flgK = 0 (residual, tangent_homo)
  drive_eps_sig      drive_eps_sig.f:16-45 (Newton iteration 1 of step 1)
  dcopy              a plain copy of one [9][V] field, the bandwidth yardstick
                     for hardware-counter calibration
  fftPcg iteration   FFT_nr3.f:214-360 (with --cg): the CG iteration group
                     (textbook operator: p = r + beta p, K4:p and p.q in pass
                     A, alpha, projection, x and r update in pass D) and the
                     residual check
  tangent_homo       tangent_homo.f:11-73 (with --homo): the nine CG solves
                     run together, host wall time of one call
  warm-start network (with --mlp; no counterpart in the reference): forward
                     pass and loss gradient of the BF16-MFMA network on 65536
                     random rows, host wall time per call

Prints one JSON line with the average ms and the algorithmic HBM GB/s of each
(bytes per voxel: 792 with K4, 144 without, 144 for the copy, 1080 for the
CG iteration group: read K4, p, r, x, write p, r, x).
"""
import argparse
import json
import time

import numpy as np

from . import B, DFM, FN1, MLP_IN, MLP_OUT, WORK, Solver
from .configs import WORKLOADS

BYTES_GKDF_K4 = 792.0   # read dF 72 + K4 648, write 72
BYTES_GKDF = 144.0      # read 72, write 72
BYTES_COPY = 144.0
BYTES_CG_IT = 1080.0    # read K4 648, p, r, x 72 each, write p, r, x 72 each


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c3_128")
    ap.add_argument("--reps", type=int, default=40)
    ap.add_argument("--skip-material", action="store_true")
    ap.add_argument("--cg", action="store_true", help="also time one CG solve")
    ap.add_argument("--homo", action="store_true", help="also time one tangent_homo call")
    ap.add_argument("--mlp", action="store_true", help="also time the warm-start network")
    ap.add_argument("--tune", default="", help="FFT launch shapes a,a0,b,c (default: untuned)")
    ap.add_argument("--f2disp", action="store_true",
                    help="also time the displacement reconstruction (host copy included)")
    args = ap.parse_args()
    wl = WORKLOADS[args.workload]
    V = wl.N ** 3
    s = Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=wl.gamma_mode)
    if args.tune:
        s.set_tune([int(v) for v in args.tune.split(",")])
    print(f"[opbench] {wl.name} N={wl.N} reps={args.reps} shapes {s.tune()}", flush=True)
    # K4 of the elastic state; a smooth strain field as the operand
    s.drive_eps_sig(1, 0)
    rng = np.random.default_rng(1)
    s.upload(DFM, 1e-4 * rng.standard_normal((9, V)))
    for _ in range(2):  # warm-up, not timed
        s.gkdf(DFM, B, 1)
        s.gkdf(DFM, B, 0)
        s.dcopy(DFM, WORK)
    s.sync()
    s.set_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        s.gkdf(DFM, B, 1)
    for _ in range(args.reps):
        s.gkdf(DFM, B, 0)
    s.sync()
    wall = time.perf_counter() - t0
    avg, n = s.kernel_times()
    gk, gk0 = avg[0], avg[3]
    # copy timing: host wall clock around a synchronised batch
    t0 = time.perf_counter()
    for _ in range(args.reps):
        s.dcopy(DFM, WORK)
    s.sync()
    cp = (time.perf_counter() - t0) / args.reps * 1e3
    res = {
        "workload": wl.name, "N": wl.N, "reps": args.reps, "shapes": s.tune(),
        "gkdf_k4_ms": gk, "gkdf_k4_GBs": BYTES_GKDF_K4 * V / gk * 1e-6,
        "gkdf_ms": gk0, "gkdf_GBs": BYTES_GKDF * V / gk0 * 1e-6,
        "dcopy_ms_wall": cp, "dcopy_GBs_wall": BYTES_COPY * V / cp * 1e-6,
        "gkdf_wall_s": wall,
    }
    if args.cg:
        # right-hand side K4:dF projected; the solve runs to 1e-8 relative
        s.gkdf(DFM, B, 1)
        s.sync()
        s.set_timing(True)
        rc, it, rel = s.pcg(B, WORK, 1e-8)
        s.sync()
        avg, n = s.kernel_times()
        res.update({"cg_rc": rc, "cg_iters": it, "cg_relres": rel,
                    "cg_group_ms": avg[0], "cg_group_GBs": BYTES_CG_IT * V / avg[0] * 1e-6,
                    "cg_check_ms": avg[2], "cg_iter_ms": avg[0] + avg[2]})
    if args.homo:
        c0 = s.counters()
        t0 = time.perf_counter()
        s.tangent_homo(1e-8)
        s.sync()
        res["homo_wall_s"] = time.perf_counter() - t0
        res["homo_cg_iterations"] = int(s.counters()[1] - c0[1])
        res["homo_ms_per_rhs_iteration"] = res["homo_wall_s"] * 1e3 / max(1, res["homo_cg_iterations"])
    if args.mlp:
        rows = 65536
        s.mlp_init(seed=1, capacity=1, batch=rows)
        rng = np.random.default_rng(2)
        X = rng.standard_normal((rows, MLP_IN)).astype(np.float32)
        T = rng.standard_normal((rows, MLP_OUT)).astype(np.float32)
        s.mlp_grad(X, T)
        s.sync()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            s.mlp_grad(X, T)
        s.sync()
        res["mlp_grad_ms_wall"] = (time.perf_counter() - t0) * 1e3 / args.reps
        res["mlp_rows"] = rows
    if args.f2disp:
        s.upload(FN1, np.eye(3).reshape(9, 1) + 1e-3 * rng.standard_normal((9, V)))
        s.f2disp(FN1)  # first call builds the cosine table
        t0 = time.perf_counter()
        for _ in range(args.reps):
            s.f2disp(FN1)
        res["f2disp_ms_wall"] = (time.perf_counter() - t0) * 1e3 / args.reps
        res["f2disp_nodes"] = (wl.N + 1) ** 3
    if not args.skip_material:
        s.upload(FN1, np.tile(np.array([1.0005, 0, 0, 0, 0.99985, 0, 0, 0, 0.99985])[:, None],
                              (1, V)))
        s.set_timing(True)
        for _ in range(3):
            s.drive_eps_sig(1, 1)
        s.sync()
        avg, n = s.kernel_times()
        res["drive_eps_sig_ms"] = avg[1]
        res["voxel_updates_per_s"] = V / (avg[1] * 1e-3)
        if args.mlp:
            # the warm start's inference path (inputs, then the fused network)
            # on every voxel, from the rolled-over plastic state; the wall time
            # includes the host copies of inputs and outputs
            s.update()
            s.upload(FN1, np.tile(np.array([1.001, 0, 0, 0, 0.9997, 0, 0, 0, 0.9997])[:, None],
                                  (1, V)))
            s.mlp_predict(2)
            t0 = time.perf_counter()
            for _ in range(3):
                s.mlp_predict(2)
            res["mlp_predict_ms_wall"] = (time.perf_counter() - t0) * 1e3 / 3
    s.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
