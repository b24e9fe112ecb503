"""CPU tests of the result-file writer (cpfft_amd/results.py, the
reference's ouresult.f flat text files) and of the deck commands that drive
it.

* The Ew.d formatter against the Fortran compiler's own output of the same
  values under format(e15.6) (amdflang, built here at test time): values
  across magnitudes and signs, zero, rounding carries (0.9999999 ->
  0.100000E+01) and random values, character for character.
* File names (wnd/wee/wes + i5.5 step + _text) and the header lines of
  ouddpa_flat_header (ouresult.f:336-397).
* The example deck's `project test` and `output results steps 2-10 by 2`.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cpfft_amd import results as R
from cpfft_amd.deck import DeckError, read_deck

HERE = os.path.dirname(os.path.abspath(__file__))


def _fortran_e15_6(values, tmp_path):
    fc = shutil.which("amdflang") or shutil.which("gfortran") or shutil.which("flang")
    if fc is None:
        pytest.skip("no Fortran compiler")
    src = tmp_path / "fe.f90"
    dat = tmp_path / "v.txt"
    np.savetxt(dat, values, fmt="%.17e")
    src.write_text(f"""program fe
  implicit none
  real(8) :: v
  integer :: i, u
  open(newunit=u, file='{dat}', status='old')
  do i = 1, {len(values)}
    read(u, *) v
    write(*, '(e15.6)') v
  end do
  close(u)
end program
""")
    exe = tmp_path / "fe"
    subprocess.run([fc, str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return out.splitlines()


def test_e15_6_matches_the_fortran_compiler(tmp_path):
    rng = np.random.default_rng(3)
    vals = [0.0, 1.0, -1.0, 123.4567, -1.23456789e-4, 999.99995, 1e-30, 6.02214076e23,
            -2.5e-7, 0.1, 0.9999999, 205.0, 5e-31 * 3, 1.0e99 * 0.5]
    vals += list(rng.standard_normal(100) * 10.0 ** rng.integers(-25, 25, 100))
    want = _fortran_e15_6(vals, tmp_path)
    got = [R.fortran_e(v) for v in vals]
    assert got == want


def test_file_names_and_header():
    assert R.file_name("wnd", 2) == "wnd00002_text"
    assert R.file_name("wes", 100001) == "wes00002_text"
    h = R._header(3, "stresses", "test", 512, 343, 10).splitlines()
    assert h[0] == "#" and h[-1] == "#"
    assert h[1] == "#  WARP3D element results: stresses       "
    assert h[2] == "#  Structure name: test    "
    assert h[3] == "#  Model nodes, elements:      512     343"
    assert h[4].startswith("#  ") and len(h[4]) == 27
    assert h[5] == "#  Load(time) step:       10"
    assert R._header(1, "displacements", "x", 8, 1, 1).splitlines()[1] == \
        "#  WARP3D nodal results: displacements  "


def test_deck_output_commands(tmp_path):
    d = read_deck(os.path.join(HERE, "golden", "reference_case", "test_mm10.in"))
    assert d.stname == "test"
    assert d.out_steps == [2, 4, 6, 8, 10]
    src = open(os.path.join(HERE, "golden", "reference_case", "test_mm10.in")).read()
    bad = tmp_path / "bad.in"
    bad.write_text(src.replace("output results steps 2-10 by 2", "output results 2-10"))
    with pytest.raises(DeckError):
        read_deck(str(bad))
    long_name = tmp_path / "long.in"
    long_name.write_text(src.replace("project test", "project name abcdefghijk"))
    assert read_deck(str(long_name)).stname == "abcdefgh"
