"""Writes oracle fixtures for the GPU tests of the larger workloads, so the
tests check full-size runs without minutes of CPU work at test time:

  c2_64_oracle.npz   the first two load steps of the 64^3, 64-grain FCC
                     workload (configs c2_64: uniaxial tension under mixed
                     boundary conditions): the inputs, the volume-averaged P
                     and the Newton / CG / mixed-BC / update counts per
                     step, and the final P and material state of every
                     256th voxel
  c4_n32_oracle.npz  the first two load steps of every 32^3 member of the
                     batched workload (configs c4_batch: its own grain
                     count, loading direction and strain rate, every
                     component of F prescribed): per member its index in
                     the batch, P per step and the counts per step
  c3_128_oracle.npz  the first two load steps of the 128^3, 256-grain FCC
                     workload (configs c3_128, the benchmark's): the inputs,
                     P_bar and the counts per step, and the final P and
                     material state of every 2048th voxel
  c5_cyclic_32_oracle.npz  config 5's crystal (BCC24) and strain increment
                     (2.5e-4 per step) in a +-0.15% tension-compression
                     cycle (about 2.5 times the yield strain), so that the
                     first reversal comes after step 6, on a 32^3, 16-grain
                     RVE, load steps 1 to 12 (through the reversal and
                     back to zero strain): P_bar and the counts per step
                     and the return code, with the workload's parameters.
                     The reference algorithm stops at step 7, the first
                     after the reversal: its Newton iteration diverges and
                     a CG solve then exhausts its 1000 iterations (rc -2)
  c4_n48_n96_oracle.npz  the first load step of the first 48^3 and the
                     first 96^3 member (the radix-3 transforms), with P
                     and the material state of every 256th voxel

The oracle is oracle/ (the CPU restatement of FFT_nr3 and the mm10
update).  Run from the repository root after building the oracle:
    python tests/golden/make_workload_oracle.py [c2|c3|c4|c4big|c5small]
(c3 took 1 h 45 min on 8 host cores, c5small 8 min).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from cpfft_amd.configs import BATCH_WORKLOADS, WORKLOADS  # noqa: E402

NSTEP = 2


def run(wl, nstep=NSTEP, allow_fail=False):
    BC, isNBC = wl.bc()
    oc = O.crystal(slip_set=wl.material.slip_set, jac_exact=wl.material.jac_exact)
    ang = wl.angles()
    P, cnt, rc, st = O.fft_nr3(oc, wl.N, ang, wl.dt, BC[:nstep], isNBC, wl.tolPCG, wl.tolNR,
                               wl.maxIter, gamma_mode=wl.gamma_mode)
    if allow_fail:
        return BC[:nstep], isNBC, ang, P, cnt, st, rc
    assert rc == 0, (wl.name, rc)
    return BC[:nstep], isNBC, ang, P, cnt, st


def c2():
    wl = WORKLOADS["c2_64"]
    BC, isNBC, ang, P, cnt, st = run(wl)
    idx = np.arange(0, wl.N ** 3, 256)
    np.savez_compressed(os.path.join(HERE, "c2_64_oracle.npz"), BC=BC, isNBC=isNBC, Pbar=P,
                        counters=cnt, idx=idx, P=st["Pn"][:, idx], state=st["sn"][:, idx])
    print("c2_64 P_bar_xx:", P[:, 0].tolist(), "counters:", cnt.tolist())


def c3():
    wl = WORKLOADS["c3_128"]
    BC, isNBC, ang, P, cnt, st = run(wl)
    idx = np.arange(0, wl.N ** 3, 2048)
    np.savez_compressed(os.path.join(HERE, "c3_128_oracle.npz"), BC=BC, isNBC=isNBC, Pbar=P,
                        counters=cnt, idx=idx, P=st["Pn"][:, idx], state=st["sn"][:, idx])
    print("c3_128 P_bar_xx:", P[:, 0].tolist(), "counters:", cnt.tolist())


C5_CYCLIC_32 = dict(N=32, ngrains=16, seed=20261202 + 55, nstep=72, strain=0.0015, cycles=3,
                    slip_set=80, run_steps=12)


def c5_cyclic_workload(p):
    """the Workload a c5_cyclic_32 parameter set describes"""
    from cpfft_amd.configs import Material, Workload
    return Workload("c5_cyclic_32", int(p["N"]), int(p["ngrains"]), int(p["seed"]),
                    int(p["nstep"]), float(p["strain"]), cycles=int(p["cycles"]),
                    material=Material(slip_set=int(p["slip_set"])))


def c5small():
    p = C5_CYCLIC_32
    wl = c5_cyclic_workload(p)
    # the run stops where the reference algorithm stops: rc < 0 and the rows
    # of the failing step and after left zero
    BC, isNBC, ang, P, cnt, st, rc = run(wl, nstep=p["run_steps"], allow_fail=True)
    np.savez_compressed(os.path.join(HERE, "c5_cyclic_32_oracle.npz"), BC=BC, isNBC=isNBC,
                        Pbar=P, counters=cnt, rc=np.array(rc),
                        **{k: np.array(v) for k, v in p.items()})
    print("c5_cyclic_32 rc", rc, "P_bar_xx:", P[:, 0].tolist(), "counters:", cnt.tolist())


def c4():
    members = BATCH_WORKLOADS["c4_batch"].members()
    keep = [k for k, m in enumerate(members) if m.N == 32]
    Ps, cs = [], []
    for k in keep:
        _, _, _, P, cnt, _ = run(members[k])
        Ps.append(P)
        cs.append(cnt)
        print(f"c4 member {k}: P_bar_xx {P[:, 0].tolist()} counters {cnt.tolist()}", flush=True)
    np.savez_compressed(os.path.join(HERE, "c4_n32_oracle.npz"), members=np.array(keep),
                        Pbar=np.array(Ps), counters=np.array(cs))


def c4big():
    members = BATCH_WORKLOADS["c4_batch"].members()
    keep = [next(k for k, m in enumerate(members) if m.N == n) for n in (48, 96)]
    out = {"members": np.array(keep)}
    for k in keep:
        wl = members[k]
        _, _, _, P, cnt, st = run(wl, nstep=1)
        idx = np.arange(0, wl.N ** 3, 256)
        tag = f"n{wl.N}"
        out.update({f"{tag}_Pbar": P, f"{tag}_counters": cnt, f"{tag}_idx": idx,
                    f"{tag}_P": st["Pn"][:, idx], f"{tag}_state": st["sn"][:, idx]})
        print(f"c4 member {k} (N={wl.N}): P_bar_xx {P[:, 0].tolist()} counters {cnt.tolist()}",
              flush=True)
    np.savez_compressed(os.path.join(HERE, "c4_n48_n96_oracle.npz"), **out)


if __name__ == "__main__":
    for w in sys.argv[1:] or ["c2", "c4", "c4big"]:
        {"c2": c2, "c3": c3, "c4": c4, "c4big": c4big, "c5small": c5small}[w]()
