"""GPU tests of the result files (ouresult.f) and of the accumulated
element strain behind them.

* The reference's own example deck run by `run_deck`, one load step per
  driver call, writing the files of `output results steps 2-10 by 2`: the
  averaged stress and counts are those of the oracle's run of the deck (as
  in test_gpu_reference_case), every file has the header and the (N+1)^3
  node or N^3 element rows of its kind, the stress file of step 10 holds
  the oracle's final per-voxel stress to the six digits written, and every
  value written is the solver's own field formatted.
* CPFFT_EPS is the sum over the rolled-over steps of each step's strain
  increment (state words CPFFT_W_D), added in step order as eps_n + uddt
  (rstgp1.f:1388-1393): bit for bit against the same sum on the host.
"""
import os

import numpy as np
import pytest

import cpfft_amd as C
from cpfft_amd import results as R
from cpfft_amd.deck import read_deck

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DECK = os.path.join(HERE, "golden", "reference_case", "test_mm10.in")


def _read(path):
    with open(path) as f:
        lines = f.read().splitlines()
    head = lines[:7]
    rows = [[float(ln[i:i + 15]) for i in range(0, len(ln), 15)] for ln in lines[7:]]
    return head, lines[7:], np.array(rows)


def test_reference_deck_result_files(tmp_path):
    g = np.load(os.path.join(HERE, "golden", "reference_case_oracle.npz"))
    keep = []
    rc, Pbar, cnt, paths = R.run_deck(DECK, str(tmp_path), gamma_mode=0, solver_out=keep)
    s = keep[0]
    print(f"\n[results] rc={rc} files {[os.path.basename(p) for p in paths]}", flush=True)
    assert rc == 0
    assert abs(Pbar - g["Pbar"]).max() <= 1e-9 * abs(g["Pbar"]).max()
    go = g["counters"]
    assert np.array_equal(cnt[:, [0, 2, 3]], go[:, [0, 2, 3]])
    assert (cnt[:, 1] <= go[:, 1]).all()
    names = sorted(os.path.basename(p) for p in paths)
    assert names == sorted(f"{k}{st:05d}_text" for k in ("wnd", "wee", "wes")
                           for st in (2, 4, 6, 8, 10))
    N, V = 7, 343
    for st in (2, 4, 6, 8, 10):
        for k, nrow, ncol, what in (("wnd", (N + 1) ** 3, 3, "#  WARP3D nodal results: displacements"),
                                    ("wee", V, 22, "#  WARP3D element results: strains"),
                                    ("wes", V, 26, "#  WARP3D element results: stresses")):
            head, lines, vals = _read(tmp_path / f"{k}{st:05d}_text")
            assert head[1].startswith(what) and head[2] == "#  Structure name: test    "
            assert head[3] == f"#  Model nodes, elements: {(N + 1) ** 3:8d}{V:8d}"
            assert head[5] == f"#  Load(time) step: {st:8d}"
            assert vals.shape == (nrow, ncol)
            if k != "wnd":
                assert not vals[:, 6:].any()
    # step 10's files are the solver's final fields, formatted
    _, lines, stress = _read(tmp_path / "wes00010_text")
    sig = s.download(C.STATE_N)[0:6].T
    assert lines[0] == "".join(R.fortran_e(v) for v in list(sig[0]) + [0.0] * 20)
    # six significant digits: |written - value| <= 5e-7 x 10^e <= 5e-6 |value|
    assert np.allclose(stress[:, :6], sig, rtol=5e-6, atol=0)
    so = g["state"][0:6].T
    assert np.allclose(stress[:, :6], so, rtol=5e-6, atol=1e-12 * abs(so).max())
    _, _, strain = _read(tmp_path / "wee00010_text")
    assert np.allclose(strain[:, :6], s.download(C.EPS).T, rtol=5e-6, atol=0)
    _, _, u = _read(tmp_path / "wnd00010_text")
    d = read_deck(DECK)
    uu = s.f2disp(lengths=d.sizes)
    uu[np.abs(uu) < 1e-30] = 0.0
    assert np.allclose(u, uu, rtol=5e-6, atol=0)
    s.close()


def test_eps_is_the_sum_of_the_steps_strain_increments():
    wl = C.Workload("eps", 8, 4, 20261202 + 41, 3, 7.5e-4)
    BC, isNBC = wl.bc()
    s = C.Solver(wl.N, wl.angles(), wl.material, dt=wl.dt, gamma_mode=1)
    assert not s.download(C.EPS).any()
    eps = np.zeros((6, wl.N ** 3))
    for k in range(3):
        rc, _, _, _ = s.compute(BC[k:k + 1], isNBC, wl.tolPCG, wl.tolNR, wl.maxIter,
                                first_step=k + 1)
        assert rc == 0
        eps = eps + s.download(C.STATE_N)[17:23]
        assert np.array_equal(s.download(C.EPS), eps), k
    assert abs(eps[0]).max() > 5e-4
    s.close()
