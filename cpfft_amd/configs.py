"""Workload definitions: material, microstructure and loading of the FCC
crystal-plasticity benchmarks, in the semantics of the reference's deck
(crystal / material / strains / loading / nonlinear analysis parameters).

Loading follows inlod (inlod.f:57-63): BC_all(:,step) is the cumulative sum
of FP_max * mults(step) plus the identity on prescribed-F diagonal terms.
Uniaxial tension along x: F_xx prescribed, P_yy = P_zz = 0 prescribed,
every other component of F prescribed with zero increment.

Time stepping: the FCC material (rate exponent 20, Voce tau_y = 16 MPa,
theta_0 = 250 MPa, reference slip rate 1e-3/s) is loaded at 1e-3/s with a
0.25 s step (2.5e-4 strain per step).  The reference's global Newton
iteration (FFT_nr3.f:99-123: full steps, no line search) does not converge
on these microstructures with coarser steps: with 1 s steps the local solves
exhaust their sub-step cuts in step 1, and with 0.5 s steps voxels at the
onset of yield flip between plastic loading and elastic unloading on every
Newton iteration at 64^3 (a period-2 cycle of the correction, whose norm
then stays fixed at ~2.3e-6 above tolNR = 1e-6).
"""
from dataclasses import dataclass, field

import numpy as np

from .microstructure import polycrystal

# FCC cubic elasticity C11/C12/C44 = 168.4/121.4/75.4 GPa expressed through the
# (e, nu, mu) inputs of elastic_type cubic (mod_crystals.f:1833-1869)
C11, C12, C44 = 168400.0, 121400.0, 75400.0
E_CUBIC = (C11 - C12) * (C11 + 2 * C12) / (C11 + C12)
NU_CUBIC = C12 / (C11 + C12)


@dataclass
class Material:
    slip_set: int = 1           # 1 fcc, 2 bcc, 6 roters, 7 bcc12, 8 bcc48 (the reference's
                                # slip_type); 80 bcc24 (bcc48's {110} and {112} systems)
    e: float = E_CUBIC
    nu: float = NU_CUBIC
    mu: float = C44
    rate_n: float = 20.0
    tau_y: float = 16.0
    tau_v: float = 132.0
    theta_0: float = 250.0
    voche_m: float = 1.0
    gamma_bar: float = 1.0e-3
    alter_mode: int = 1
    atol: float = 1.0e-5
    atol1: float = 1.0e-5
    rtol: float = 5.0e-5
    rtol1: float = 1.0e-5
    miter: int = 30
    jac_exact: int = 1


@dataclass
class Workload:
    name: str
    N: int
    ngrains: int
    seed: int
    nstep: int
    strain: float               # final F_xx - 1
    dt: float = 0.25
    tolPCG: float = 1.0e-8
    tolNR: float = 1.0e-6
    maxIter: int = 25
    gamma_mode: int = 1         # 1 textbook projection, 0 reference operator
    material: Material = field(default_factory=Material)
    cycles: int = 0             # > 0: tension-compression cycles of amplitude `strain`
    direction: tuple = None     # unit n: every F component prescribed, F = I + eps n (x) n

    def angles(self):
        return polycrystal(self.N, self.ngrains, self.seed)[1]

    def bc(self):
        """(BC_all [nstep][9], isNBC [9]) for uniaxial loading along x:
        monotonic tension to `strain`, or with cycles > 0 a triangle wave
        0 -> +strain -> -strain -> 0 repeated `cycles` times (mults of
        either sign, as a loading deck may give them).  With `direction`
        set, uniaxial strain along n instead: every component of F
        prescribed (no mixed boundary condition)."""
        isNBC = np.zeros(9, np.int32)
        FP_max = np.zeros(9)
        if self.direction is not None:
            # F_bar = I + eps n (x) n, all nine components prescribed
            n = np.asarray(self.direction, dtype=np.float64)
            FP_max[:] = self.strain * np.outer(n, n).ravel()
        else:
            isNBC[[4, 8]] = 1
            FP_max[0] = self.strain
        if self.cycles > 0:
            q = self.nstep // (4 * self.cycles)
            assert q * 4 * self.cycles == self.nstep, "nstep must be a multiple of 4 * cycles"
            quarter = np.full(q, 1.0 / q)
            mults = np.tile(np.concatenate([quarter, -quarter, -quarter, quarter]), self.cycles)
        else:
            mults = np.full(self.nstep, 1.0 / self.nstep)
        BC = np.cumsum(FP_max[None, :] * mults[:, None], axis=0)
        for d in (0, 4, 8):
            if not isNBC[d]:
                BC[:, d] += 1.0
        return np.ascontiguousarray(BC), isNBC


# seeds: numpy default_rng(20261202 + k) for workload k
WORKLOADS = {
    "c1_16": Workload("c1_16", 16, 8, 20261202 + 1, 80, 0.02),
    "c2_64": Workload("c2_64", 64, 64, 20261202 + 2, 160, 0.04),
    "c3_128": Workload("c3_128", 128, 256, 20261202 + 3, 240, 0.06),
    # BCC (the first 24 systems of bcc48), cyclic +-1.5%, 3 cycles; the cubic
    # elastic constants of the FCC material are kept
    "c5_256": Workload("c5_256", 256, 1024, 20261202 + 5, 720, 0.015, cycles=3,
                       material=Material(slip_set=80)),
}


@dataclass
class BatchWorkload:
    """Many independent RVEs of uneven size solved together: grid edges
    drawn from `edges`, grain counts uniform in [grains[0], grains[1]],
    loading F_bar = I + eps n (x) n along a direction n uniform on the
    sphere, at a strain rate log-uniform in [rates[0], rates[1]] per second
    with `dstrain` strain per step (dt = dstrain / rate)."""
    name: str
    nrve: int
    edges: tuple
    grains: tuple
    rates: tuple
    nstep: int
    dstrain: float
    seed: int

    def members(self):
        rng = np.random.default_rng(self.seed)
        N = rng.choice(np.asarray(self.edges), self.nrve)
        ng = rng.integers(self.grains[0], self.grains[1] + 1, self.nrve)
        d = rng.standard_normal((self.nrve, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rate = 10.0 ** rng.uniform(np.log10(self.rates[0]), np.log10(self.rates[1]), self.nrve)
        return [Workload(f"{self.name}_{k}", int(N[k]), int(ng[k]), self.seed * 100 + k,
                         self.nstep, self.nstep * self.dstrain, dt=self.dstrain / float(rate[k]),
                         direction=tuple(d[k])) for k in range(self.nrve)]


# 24 RVEs with edges 32-96, 16-128 grains, each with its own loading
# direction and strain rate in [1e-4, 1e-2]/s, 30 steps of 2.5e-4 strain
BATCH_WORKLOADS = {
    "c4_batch": BatchWorkload("c4_batch", 24, (32, 48, 64, 96), (16, 128), (1e-4, 1e-2), 30,
                              2.5e-4, 20261202 + 4),
}
