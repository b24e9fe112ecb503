"""CPU tests of the workload definitions: boundary values in the semantics
of inlod (inlod.f:57-63: cumulative FP_max * mults, identity added on
prescribed-F diagonal terms) and the batched uneven-RVE generator."""
import numpy as np

from cpfft_amd.configs import BATCH_WORKLOADS, WORKLOADS, Workload


def test_uniaxial_tension_mixed_bc():
    wl = WORKLOADS["c1_16"]
    BC, isNBC = wl.bc()
    assert isNBC.tolist() == [0, 0, 0, 0, 1, 0, 0, 0, 1]
    eps = np.arange(1, wl.nstep + 1) * wl.strain / wl.nstep
    assert np.allclose(BC[:, 0], 1.0 + eps, rtol=0, atol=1e-15)
    # prescribed-P components carry the target stress (zero); the other
    # prescribed-F components stay at the identity
    assert not BC[:, [4, 8]].any()
    assert not BC[:, [1, 2, 3, 5, 6, 7]].any()


def test_cyclic_loading_returns_to_zero():
    wl = WORKLOADS["c5_256"]
    BC, _ = wl.bc()
    e = BC[:, 0] - 1.0
    assert abs(e.max() - wl.strain) < 1e-14 and abs(e.min() + wl.strain) < 1e-14
    assert abs(e[-1]) < 1e-14
    # three cycles: the strain crosses zero going up at the start of each
    ups = np.sum((e[:-1] < 0) & (e[1:] >= 0))
    assert ups == wl.cycles - 1 or ups == wl.cycles


def test_uniaxial_strain_along_a_direction():
    n = np.array([1.0, 2.0, -2.0]) / 3.0
    wl = Workload("d", 8, 2, 1, 5, 5 * 2.5e-4, direction=tuple(n))
    BC, isNBC = wl.bc()
    assert not isNBC.any()
    for k in range(5):
        F = np.eye(3) + (k + 1) * 2.5e-4 * np.outer(n, n)
        assert np.allclose(BC[k].reshape(3, 3), F, rtol=0, atol=1e-15)


def test_batch_workload_members():
    bw = BATCH_WORKLOADS["c4_batch"]
    m1, m2 = bw.members(), bw.members()
    assert len(m1) == bw.nrve == 24
    assert [(w.N, w.ngrains, w.dt, w.direction) for w in m1] == \
        [(w.N, w.ngrains, w.dt, w.direction) for w in m2]
    assert {w.N for w in m1} <= set(bw.edges) and len({w.N for w in m1}) > 1
    for w in m1:
        assert bw.grains[0] <= w.ngrains <= bw.grains[1]
        rate = bw.dstrain / w.dt
        assert bw.rates[0] <= rate <= bw.rates[1]
        assert abs(np.linalg.norm(w.direction) - 1.0) < 1e-14
        assert w.nstep == bw.nstep and abs(w.strain - bw.nstep * bw.dstrain) < 1e-15
