"""The reference's own crystal-plasticity example as a fixture.

Copies the two input files the reference ships for it,
examples/test_mm10.in (the deck: a 7^3 grid, one bcc48 crystal type with
isotropic elasticity and Voce hardening, every component of F prescribed,
10 load steps of 10 s) and examples/angle_bc.in (the per-element Kocks
angles it reads: a bicrystal, 196 elements at (0, 0, 0) and 147 at
(45, 0, 0)), into tests/golden/reference_case/, and writes
reference_case_oracle.npz: the oracle's whole run of that deck with the
reference's operator (gamma_mode 0; N = 7 is odd, where it is the
textbook projection) and the reference's local Jacobian (jac_exact 0) --
the volume-averaged P per step, the Newton / CG / mixed-BC / update
counts per step, and the final P and material state of every voxel.
Run in the container that holds the reference, after building the oracle:
    python tests/golden/make_reference_case.py
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from cpfft_amd.deck import read_deck  # noqa: E402

CASE = os.path.join(HERE, "reference_case")
SRC = "/root/reference/examples"


def oracle_crystal(m):
    return O.crystal(slip_set=m.slip_set, e=m.e, nu=m.nu, mu=m.mu, rate_n=m.rate_n,
                     tau_y=m.tau_y, tau_v=m.tau_v, theta_0=m.theta_0, voche_m=m.voche_m,
                     alter_mode=bool(m.alter_mode), gamma_bar=m.gamma_bar,
                     jac_exact=m.jac_exact)


def run(deck_path):
    d = read_deck(deck_path)
    m = d.material()
    BC, isNBC = d.bc()
    ang = d.angles()
    return d, BC, isNBC, ang, O.fft_nr3(oracle_crystal(m), d.N, ang, d.dt, BC, isNBC,
                                        d.tolPCG, d.tolNR, d.maxIter, gamma_mode=0)


def main():
    os.makedirs(CASE, exist_ok=True)
    for f in ("test_mm10.in", "angle_bc.in"):
        shutil.copyfile(os.path.join(SRC, f), os.path.join(CASE, f))
    d, BC, isNBC, ang, (Pbar, cnt, rc, st) = run(os.path.join(CASE, "test_mm10.in"))
    assert rc == 0, rc
    np.savez_compressed(os.path.join(HERE, "reference_case_oracle.npz"), BC=BC, isNBC=isNBC,
                        angles=ang, Pbar=Pbar, counters=cnt, P=st["Pn"], state=st["sn"],
                        N=d.N, dt=d.dt)
    print("P_bar_xx per step:", Pbar[:, 0].round(4).tolist())


if __name__ == "__main__":
    main()
