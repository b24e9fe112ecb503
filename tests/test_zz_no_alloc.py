"""cpfft_gkdf_noalloc (the context-free G_K_dF that allocates nothing)
against cpfft_gkdf through the C-ABI.  The test plays the caller: it owns
the device arrays and the stream (made with the HIP runtime the library is
linked against), and the entry point must give results bit-identical to
the context's operator (the same kernels), call after call, with later
calls on smaller grids (the shared scratch keeps the earlier call's
contents), and reject sizes above its maximum."""
import ctypes

import numpy as np
import pytest

import cpfft_amd as C
from cpfft_amd.microstructure import polycrystal

pytestmark = pytest.mark.gpu


def _hip():
    """the HIP runtime instance libcpfft_amd.so is linked against"""
    C.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


class DeviceArray:
    def __init__(self, hip, host):
        self.hip, self.n = hip, host.nbytes
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), self.n) == 0
        assert hip.hipMemcpy(self.p, host.ctypes.data, self.n, 1) == 0  # host to device

    def get(self, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.n, 2) == 0  # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _context_result(N, mode, F, K4):
    _, ang = polycrystal(N, 3, 31)
    s = C.Solver(N, ang, C.Material(), dt=0.25, gamma_mode=mode)
    s.upload(C.DFM, F)
    s.upload(C.K4, K4)
    s.gkdf(C.DFM, C.B, 0)
    g0 = s.download(C.B)
    s.gkdf(C.DFM, C.B, 1)
    g1 = s.download(C.B)
    s.close()
    return g0, g1


def test_gkdf_noalloc_matches_cpfft_gkdf():
    L = C.lib()
    hip = _hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(stream), 1) == 0  # non-blocking
    # N = 48 (the maximum, planned path), then smaller grids: 16 (planned)
    # and 7 (runtime-plan path), each call reusing the scratch
    for N, mode in ((48, 1), (16, 0), (7, 1)):
        rng = np.random.default_rng(100 + N)
        F = rng.standard_normal((9, N ** 3))
        K4 = rng.standard_normal((81, N ** 3))
        g0, g1 = _context_result(N, mode, F, K4)
        dF, dK = DeviceArray(hip, F), DeviceArray(hip, K4)
        dO = DeviceArray(hip, np.full((9, N ** 3), np.nan))
        print(f"\n[noalloc] N={N} mode={mode}", flush=True)
        rc = L.cpfft_gkdf_noalloc(N, mode, dF.p, None, dO.p, stream)
        assert rc == 0, L.cpfft_last_error().decode()
        assert hip.hipStreamSynchronize(stream) == 0
        assert np.array_equal(dO.get((9, N ** 3)), g0)
        rc = L.cpfft_gkdf_noalloc(N, mode, dF.p, dK.p, dO.p, stream)
        assert rc == 0, L.cpfft_last_error().decode()
        assert hip.hipStreamSynchronize(stream) == 0
        assert np.array_equal(dO.get((9, N ** 3)), g1)
        for a in (dF, dK, dO):
            a.free()
    # rejected before anything is launched
    dF = DeviceArray(hip, np.zeros((9, 8 ** 3)))
    dO = DeviceArray(hip, np.zeros((9, 8 ** 3)))
    assert L.cpfft_gkdf_noalloc(64, 1, dF.p, None, dO.p, stream) == -1
    assert b"CPFFT_NOALLOC_MAX_N" in L.cpfft_last_error()
    assert L.cpfft_gkdf_noalloc(1, 1, dF.p, None, dO.p, stream) == -1
    assert L.cpfft_gkdf_noalloc(8, 2, dF.p, None, dO.p, stream) == -1
    assert L.cpfft_gkdf_noalloc(8, 1, dF.p, None, dF.p, stream) == -1
    assert L.cpfft_gkdf_noalloc(8, 1, None, None, dO.p, stream) == -1
    assert L.cpfft_gkdf_noalloc(17, 1, dF.p, None, dO.p, stream) == -3
    dF.free()
    dO.free()
    assert hip.hipStreamDestroy(stream) == 0
