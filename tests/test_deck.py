"""CPU tests of the deck and angle-file reader (cpfft_amd/deck.py) on the
reference's own example deck (tests/golden/reference_case, copied from
examples/) and on synthetic decks, and of the reference-case fixture
against the oracle."""
import os

import numpy as np
import pytest

import oracle_lib as O
from cpfft_amd.configs import WORKLOADS
from cpfft_amd.deck import DeckError, read_angle_file, read_deck

HERE = os.path.dirname(os.path.abspath(__file__))
CASE = os.path.join(HERE, "golden", "reference_case")


def test_reference_example_deck():
    d = read_deck(os.path.join(CASE, "test_mm10.in"))
    assert d.N == 7 and d.computed
    cr = d.crystals[1]
    assert cr.slip_type == 8 and cr.elastic_type == 1 and cr.hardening == "voce"
    assert (cr.e, cr.nu, cr.theta_0, cr.voche_m, cr.tau_v, cr.tau_y) == (
        200000.0, 0.3, 0.01, 1.0, 5000.0, 205.0)
    assert cr.gamma_bar == 4.0e-05 and cr.alter_mode
    # `harden_n 5,48`: the comma makes the crystal reader take the next
    # line (incrystal.f:987-988), so 48 is never read
    assert cr.harden_n == 5.0
    m = d.material()
    assert m.slip_set == 8 and m.rate_n == 5.0 and m.jac_exact == 0
    assert abs(m.mu - 200000.0 / 2.6) < 1e-9
    assert d.sizes == (100.0, 100.0, 100.0)
    assert (d.maxIter, d.tolNR, d.tolPCG, d.dt) == (10, 1e-5, 1e-10, 10.0)
    BC, isNBC = d.bc()
    assert not isNBC.any() and BC.shape == (10, 9)
    # inlod.f:57-63: running sums of FP_max * 0.1 plus the identity
    k = np.arange(1, 11)
    assert np.allclose(BC[:, 0], 1 + 0.003 * k, rtol=0, atol=1e-15)
    assert np.allclose(BC[:, 4], 1 - 0.001 * k, rtol=0, atol=1e-15)
    assert np.allclose(BC[:, 8], BC[:, 4], rtol=0, atol=0)
    assert not BC[:, [1, 2, 3, 5, 6, 7]].any()
    ang = d.angles()
    assert ang.shape == (3, 343)
    u, cnt = np.unique(ang.T, axis=0, return_counts=True)
    assert u.tolist() == [[0, 0, 0], [45, 0, 0]] and cnt.tolist() == [196, 147]


def test_angle_file_is_read_forward_by_element(tmp_path):
    """read_defs (mod_crystals.f:2246-2300) reads forward to each element's
    record, skipping others; a missing element is an error."""
    p = tmp_path / "a.in"
    p.write_text("1, 10, 20, 30\n7, 0, 0, 0\n2 11 21 31\n3,12.5,22,32\n")
    a = read_angle_file(str(p), 3)
    assert a.T.tolist() == [[10, 20, 30], [11, 21, 31], [12.5, 22, 32]]
    with pytest.raises(DeckError):
        read_angle_file(str(p), 4)


def test_mixed_bc_deck_matches_the_workload_definition(tmp_path):
    """A deck for the 16^3 workload's loading (F_xx prescribed, P_yy = P_zz
    = 0, 80 equal steps) gives the BC_all the workload builds."""
    wl = WORKLOADS["c1_16"]
    (tmp_path / "ang.in").write_text(
        "".join(f"{e + 1}, 10.0, 20.0, 30.0\n" for e in range(8 ** 3)))
    deck = f"""c a synthetic deck
number of grid 8
crystal 1
      properties slip_type fcc elastic_type cubic,
      hardening voce e {wl.material.e!r} nu {wl.material.nu!r} mu 75400,
      harden_n 20 alter_mode on theta_0 250 tau_y 16 tau_v 132,
      gamma_bar 1.0e-3
material m
      properties cp angle_convention kocks n_crystals 1 angle_type degrees,
      crystal_input single crystal_type 1 orientation_input file filename 'ang.in'
elements
1-512 material m
strains
  F_xx {wl.strain!r}
  P_yy 0.0
  P_zz 0.0
loading t
  step 1-{wl.nstep} constraints {1.0 / wl.nstep!r}
nonlinear analysis parameters
  maximum iterations 25
  convergence tolerance NR 1.0e-6 CG 1.0e-8
  time step 0.25
compute
stop
"""
    (tmp_path / "d.in").write_text(deck)
    d = read_deck(str(tmp_path / "d.in"))
    BC, isNBC = d.bc()
    BCw, isNBCw = wl.bc()
    assert np.array_equal(isNBC, isNBCw)
    assert abs(BC - BCw).max() < 1e-15
    m = d.material()
    assert (m.slip_set, m.rate_n, m.tau_y, m.tau_v, m.theta_0, m.gamma_bar) == (1, 20, 16, 132, 250, 1e-3)
    assert m.e == wl.material.e and m.nu == wl.material.nu and m.mu == 75400.0
    assert (d.angles() == np.array([[10.0], [20.0], [30.0]])).all()


def test_deck_errors(tmp_path):
    p = tmp_path / "bad.in"
    p.write_text("number of grid 4\nstop\n")
    with pytest.raises(DeckError):  # no nonlinear analysis parameters
        read_deck(str(p))
    d = read_deck(os.path.join(os.path.dirname(HERE), "tests", "golden", "reference_case",
                               "test_mm10.in"))
    d.crystals[1].hardening = "mts"
    with pytest.raises(DeckError):
        d.material()


def test_reference_case_fixture_is_the_oracle_run():
    g = np.load(os.path.join(HERE, "golden", "reference_case_oracle.npz"))
    d = read_deck(os.path.join(CASE, "test_mm10.in"))
    BC, isNBC = d.bc()
    assert np.array_equal(g["BC"], BC) and np.array_equal(g["isNBC"], isNBC)
    assert np.array_equal(g["angles"], d.angles())
    m = d.material()
    oc = O.crystal(slip_set=m.slip_set, e=m.e, nu=m.nu, mu=m.mu, rate_n=m.rate_n,
                   tau_y=m.tau_y, tau_v=m.tau_v, theta_0=m.theta_0, voche_m=m.voche_m,
                   alter_mode=bool(m.alter_mode), gamma_bar=m.gamma_bar, jac_exact=0)
    P, c, rc, st = O.fft_nr3(oc, d.N, d.angles(), d.dt, BC, isNBC, d.tolPCG, d.tolNR,
                             d.maxIter, gamma_mode=0)
    assert rc == 0
    assert np.array_equal(P, g["Pbar"]) and np.array_equal(c, g["counters"])
    assert np.array_equal(st["Pn"], g["P"]) and np.array_equal(st["sn"], g["state"])
    # every component of F prescribed: mean F is the prescription and P
    # rises with the tension
    assert (np.diff(P[:, 0]) > 0).all()


@pytest.mark.parametrize("name,slip_set", [("fcc", 1), ("bcc", 2), ("roters", 6), ("bcc12", 7),
                                           ("bcc48", 8)])
def test_slip_type_names(tmp_path, name, slip_set):
    """incrystal.f:57-84's slip_type names map to the library's slip sets;
    the ones it does not implement are refused."""
    src = open(os.path.join(CASE, "test_mm10.in")).read().replace("slip_type bcc48", f"slip_type {name}")
    p = tmp_path / "d.in"
    p.write_text(src)
    assert read_deck(str(p)).material().slip_set == slip_set


def test_unimplemented_slip_type_is_refused(tmp_path):
    src = open(os.path.join(CASE, "test_mm10.in")).read().replace("slip_type bcc48", "slip_type hcp6")
    p = tmp_path / "d.in"
    p.write_text(src)
    with pytest.raises(DeckError):
        read_deck(str(p)).material()
