"""Writes tests/golden/c1_16_oracle.npz: the oracle's (oracle/, the CPU
restatement of FFT_nr3 and the mm10 update) whole load-step run of the
16^3, 8-grain FCC workload (configs c1_16: uniaxial tension under mixed
boundary conditions, 80 steps of 0.25 s to 2%), so the GPU test can check
the full run without spending minutes of CPU on the oracle.

Contents: the inputs (boundary values, mask, Euler angles), the oracle's
volume-averaged P per step, its Newton / CG / mixed-BC / update counts per
step, and the final first Piola stress and material state on every 8th
voxel.  Run from the repository root after building the oracle:
    python tests/golden/make_c1_16_oracle.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from cpfft_amd.configs import WORKLOADS  # noqa: E402


def main():
    wl = WORKLOADS["c1_16"]
    BC, isNBC = wl.bc()
    ang = wl.angles()
    oc = O.crystal(jac_exact=wl.material.jac_exact)
    Pbar, cnt, rc, st = O.fft_nr3(oc, wl.N, ang, wl.dt, BC, isNBC, wl.tolPCG, wl.tolNR,
                                  wl.maxIter, gamma_mode=wl.gamma_mode)
    assert rc == 0, rc
    idx = np.arange(0, wl.N ** 3, 8)
    np.savez_compressed(os.path.join(HERE, "c1_16_oracle.npz"), BC=BC, isNBC=isNBC, angles=ang,
                        Pbar=Pbar, counters=cnt, idx=idx, P=st["Pn"][:, idx],
                        state=st["sn"][:, idx], N=wl.N, dt=wl.dt, gamma_mode=wl.gamma_mode)
    print("P_bar_xx per step:", Pbar[:, 0].round(4).tolist())


if __name__ == "__main__":
    main()
