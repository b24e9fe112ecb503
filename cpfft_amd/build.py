"""Build the MI355X library and its Fortran driver, in-tree.

  libcpfft_amd.so     HIP kernels + C-ABI (include/cpfft.h), gfx950 only
  libcpfft_driver.so  Fortran load-step driver (cpfft_compute), amdflang, and
                      the C++ batch runner (cpfft_compute_batch)

Run `python -m cpfft_amd.build` (or __graft_entry__.build()).
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(ROOT, "_build")

HIP_SOURCES = ["cpfft_kernels.hip", "fft_pencil.hip", "mlp.hip", "f2disp.hip", "cpfft_api.hip",
               "cpfft_noalloc.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are disabled on this platform
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build_hip():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    out = os.path.join(PKG, "libcpfft_amd.so")
    tmp = out + ".tmp"
    _run([hipcc] + HIP_FLAGS + [os.path.join(CSRC, s) for s in HIP_SOURCES] + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_fortran():
    """The Fortran load-step driver and the C++ batch runner above it."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(PKG, "libcpfft_driver.so")
    tmp = out + ".tmp"
    obj = os.path.join(BUILD, "batch_runner.o")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-c",
          os.path.join(FSRC, "batch_runner.cpp"), "-o", obj])
    _run([fc, "-O2", "-fPIC", "-shared", "-module-dir", BUILD,
          os.path.join(FSRC, "cpfft_bindings.f90"), os.path.join(FSRC, "fft_nr3.f90"), obj,
          "-L" + PKG, "-lcpfft_amd", "-lstdc++", "-lpthread", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_hostcheck():
    """Test-only host build of the per-voxel update (tests/hostcheck), so the
    CPU test suite can check the kernel's algebra against the oracle."""
    src = os.path.join(ROOT, "tests", "hostcheck", "voxel_host.cpp")
    out = os.path.join(ROOT, "tests", "hostcheck", "libvoxel_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-Wno-unknown-pragmas", src, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)


def build_all():
    build_hip()
    build_fortran()
    build_oracle()
    build_hostcheck()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "fortran": build_fortran, "oracle": build_oracle,
         "hostcheck": build_hostcheck}[w]()
