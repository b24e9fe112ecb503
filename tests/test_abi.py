"""CPU checks of the drop-in boundary: the libraries load, export every
function include/*.h declares, the Python mirror and the Fortran bindings
name only declared functions, and the ctypes structures have the layout a
C compiler gives the header's structures.  No compute call is made (no GPU
is needed)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

import cpfft_amd as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z_0-9]*\s*\**\s*(cpfft_[a-z_0-9]+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


def test_core_library_exports_every_declared_function():
    names = declared("cpfft.h")
    assert len(names) >= 30
    L = C.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert sorted(C.EXPORTS) == names


def test_driver_library_exports_every_declared_function():
    names = declared("cpfft_driver.h")
    assert names == ["cpfft_compute", "cpfft_compute_batch"]
    D = C.driver()
    assert all(hasattr(D, n) for n in names)
    assert sorted(C.DRIVER_EXPORTS) == names


def test_fortran_bindings_name_declared_functions():
    src = open(os.path.join(ROOT, "cpfft_amd", "fortran", "cpfft_bindings.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(cpfft_[a-z_0-9]+)"\)', src))
    assert bound and bound <= set(declared("cpfft.h"))


LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "cpfft_driver.h"
#define F(T, m) printf(#T " " #m " %zu\n", offsetof(T, m))
int main(void) {
  printf("cpfft_material size %zu\n", sizeof(cpfft_material));
  F(cpfft_material, e); F(cpfft_material, alter_mode); F(cpfft_material, atol);
  F(cpfft_material, miter); F(cpfft_material, jac_exact);
  printf("cpfft_batch_item size %zu\n", sizeof(cpfft_batch_item));
  F(cpfft_batch_item, nstep); F(cpfft_batch_item, BC_all); F(cpfft_batch_item, tolPCG);
  F(cpfft_batch_item, maxIter); F(cpfft_batch_item, weight); F(cpfft_batch_item, Pbar_out);
  F(cpfft_batch_item, counters); F(cpfft_batch_item, times); F(cpfft_batch_item, rc);
  return 0;
}
"""


def test_struct_layouts_match_the_header():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(src, "w").write(LAYOUT_C)
        r = subprocess.run(["gcc", "-I", INC, src, "-o", exe], capture_output=True, text=True)
        if r.returncode != 0:
            pytest.fail(r.stderr)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in filter(None, lines):
        t, m, v = ln.split()
        got[(t, m)] = int(v)
    py = {"cpfft_material": C.CMaterial, "cpfft_batch_item": C.CBatchItem}
    for (t, m), v in got.items():
        S = py[t]
        want = ctypes.sizeof(S) if m == "size" else getattr(S, m).offset
        assert want == v, (t, m, want, v)
