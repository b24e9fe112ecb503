"""CPU checks of the committed golden fixtures (tests/golden/): each holds
the inputs of the workload it names, satisfies the loading it encodes, and
is what the oracle computes now (its first steps recomputed bitwise)."""
import os

import numpy as np
import pytest

import oracle_lib as O
from cpfft_amd.configs import WORKLOADS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_c1_16_fixture():
    g = np.load(os.path.join(GOLDEN, "c1_16_oracle.npz"))
    wl = WORKLOADS["c1_16"]
    BC, isNBC = wl.bc()
    assert np.array_equal(g["BC"], BC) and np.array_equal(g["isNBC"], isNBC)
    assert np.array_equal(g["angles"], wl.angles())
    Pb = g["Pbar"]
    assert Pb.shape == (wl.nstep, 9) and np.isfinite(Pb).all()
    # uniaxial tension under mixed BC: P_yy = P_zz = 0 to the mixed-BC
    # tolerance, P_xx rising
    assert abs(Pb[:, [4, 8]]).max() <= 2 * wl.tolNR * abs(Pb[:, 0]).max()
    assert (np.diff(Pb[:, 0]) > 0).all()
    # the oracle gives the same first two steps today
    oc = O.crystal(jac_exact=wl.material.jac_exact)
    P2, c2, rc, _ = O.fft_nr3(oc, wl.N, g["angles"], wl.dt, BC[:2], isNBC, wl.tolPCG, wl.tolNR,
                              wl.maxIter, gamma_mode=wl.gamma_mode)
    assert rc == 0
    assert np.array_equal(P2, Pb[:2]) and np.array_equal(c2, g["counters"][:2])


def test_workload_fixtures_name_their_workloads():
    """The larger workloads' fixtures (make_workload_oracle.py; recomputing
    them takes minutes to hours of CPU, so only their inputs are checked
    here): the 64^3 loading and the batch members they hold."""
    from cpfft_amd.configs import BATCH_WORKLOADS
    g = np.load(os.path.join(GOLDEN, "c2_64_oracle.npz"))
    wl = WORKLOADS["c2_64"]
    BC, isNBC = wl.bc()
    assert np.array_equal(g["BC"], BC[:2]) and np.array_equal(g["isNBC"], isNBC)
    assert g["Pbar"].shape == (2, 9) and g["state"].shape[1] == g["idx"].size
    members = BATCH_WORKLOADS["c4_batch"].members()
    g = np.load(os.path.join(GOLDEN, "c4_n32_oracle.npz"))
    assert [k for k, m in enumerate(members) if m.N == 32] == g["members"].tolist()
    assert g["Pbar"].shape == (len(g["members"]), 2, 9)
    g = np.load(os.path.join(GOLDEN, "c4_n48_n96_oracle.npz"))
    assert [members[k].N for k in g["members"]] == [48, 96]
    for n in (48, 96):
        assert g[f"n{n}_Pbar"].shape == (1, 9)
        assert g[f"n{n}_idx"].tolist() == list(range(0, n ** 3, 256))
        assert g[f"n{n}_state"].shape == (26 + 2 * 12, g[f"n{n}_idx"].size)


def _fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not generated (tests/golden/make_workload_oracle.py)")
    return np.load(path)


def test_c3_128_fixture_inputs():
    g = _fixture("c3_128_oracle.npz")
    wl = WORKLOADS["c3_128"]
    BC, isNBC = wl.bc()
    assert np.array_equal(g["BC"], BC[:2]) and np.array_equal(g["isNBC"], isNBC)
    assert g["Pbar"].shape == (2, 9) and g["idx"].tolist() == list(range(0, 128 ** 3, 2048))
    assert g["state"].shape == (26 + 2 * 12, g["idx"].size)


def test_c5_cyclic_32_fixture_inputs():
    """config 5's crystal and strain increment in a smaller cycle on 32^3:
    the parameters it stores give its inputs"""
    g = _fixture("c5_cyclic_32_oracle.npz")
    from cpfft_amd.configs import Material, Workload
    wl5 = WORKLOADS["c5_256"]
    wl = Workload("c5_cyclic_32", int(g["N"]), int(g["ngrains"]), int(g["seed"]),
                  int(g["nstep"]), float(g["strain"]), cycles=int(g["cycles"]),
                  material=Material(slip_set=int(g["slip_set"])))
    assert (wl.dt, wl.material) == (wl5.dt, wl5.material)

    def increment(w):
        return w.strain / (w.nstep // (4 * w.cycles))
    assert abs(increment(wl) - increment(wl5)) < 1e-15
    n = int(g["run_steps"])
    BC, isNBC = wl.bc()
    assert np.array_equal(g["BC"], BC[:n]) and np.array_equal(g["isNBC"], isNBC)
    assert g["Pbar"].shape == (n, 9) and n > wl.nstep // (4 * wl.cycles)  # past the reversal
    # the steps the oracle completed come first; it stops with rc < 0 or none
    done = int(np.count_nonzero(g["counters"].any(axis=1)))
    assert g["counters"][:done].all() and not g["counters"][done:].any()
    assert (int(g["rc"]) < 0) == (done < n)
