"""ctypes bindings to the oracle (oracle/liboracle.so).

The oracle is test infrastructure: a CPU restatement of the reference's
algorithm used only as the parity checker.  Only tests, the smoke check and
bench.py's cpu_baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

MAX_SLIP = 48
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


class OrcCrystal(ctypes.Structure):
    _fields_ = [
        ("nslip", ctypes.c_int),
        ("bi", ctypes.c_double * (MAX_SLIP * 3)),
        ("ni", ctypes.c_double * (MAX_SLIP * 3)),
        ("stiff", ctypes.c_double * 36),
        ("rate_n", ctypes.c_double),
        ("tau_y", ctypes.c_double),
        ("tau_v", ctypes.c_double),
        ("theta_0", ctypes.c_double),
        ("voche_m", ctypes.c_double),
        ("eps_dot_0_y", ctypes.c_double),
        ("iD_v", ctypes.c_double),
        ("alter_mode", ctypes.c_int),
        ("atol", ctypes.c_double),
        ("atol1", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("rtol1", ctypes.c_double),
        ("miter", ctypes.c_int),
        ("jac_exact", ctypes.c_int),
    ]


class OrcProblem(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int),
        ("V", ctypes.c_long),
        ("cp", ctypes.POINTER(OrcCrystal)),
        ("angles", _dp),
        ("dt", ctypes.c_double),
        ("tolPCG", ctypes.c_double),
        ("tolNR", ctypes.c_double),
        ("maxIter", ctypes.c_int),
        ("gamma_mode", ctypes.c_int),
        ("Fn", _dp), ("Fn1", _dp), ("Pn", _dp), ("Pn1", _dp), ("K4", _dp),
        ("sn", _dp), ("snp1", _dp),
    ]


def nwords(nslip):
    return 26 + 2 * nslip


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_crystal_init.argtypes = [ctypes.POINTER(OrcCrystal), ctypes.c_int] + \
            [ctypes.c_double] * 8 + [ctypes.c_int, ctypes.c_double]
        L.orc_voxel_update.argtypes = [ctypes.POINTER(OrcCrystal), _dp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp,
                                       _dp, _dp, _ip]
        L.orc_drive_eps_sig.argtypes = [ctypes.POINTER(OrcCrystal), ctypes.c_long, _dp,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                                        _dp, _dp, _dp, _dp, _dp, _lp]
        L.orc_state_init.argtypes = [ctypes.POINTER(OrcCrystal), ctypes.c_long, _dp]
        L.orc_rotation_matrix.argtypes = [_dp, _dp]
        L.orc_rtcmp1.argtypes = [_dp, _dp]
        L.orc_cep2A.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_double, _dp, _dp, _dp, _dp,
                                ctypes.c_double, _dp]
        L.orc_fft_create.argtypes = [ctypes.c_int]
        L.orc_fft_create.restype = ctypes.c_void_p
        L.orc_fft_create2.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_fft_create2.restype = ctypes.c_void_p
        L.orc_fft_destroy.argtypes = [ctypes.c_void_p]
        L.orc_gkdf.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_int]
        L.orc_ddot42n.argtypes = [ctypes.c_long, _dp, _dp, _dp]
        L.orc_pcg.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, _ip, _dp]
        L.orc_pcg_limit.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                    _ip, _dp]
        L.orc_tangent_homo.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.orc_nbc_update.argtypes = [_dp, _dp, _dp, _dp, _ip]
        L.orc_fft_nr3.argtypes = [ctypes.POINTER(OrcProblem), ctypes.c_int, _dp, _ip, _dp,
                                  _ip]
        _lib = L
    return _lib


def P(a):
    """pointer to a C-contiguous float64 numpy array"""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def IP(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def crystal(slip_set=1, e=None, nu=None, mu=75400.0, C11=168400.0, C12=121400.0,
            rate_n=20.0, tau_y=16.0, tau_v=132.0, theta_0=250.0, voche_m=1.0,
            alter_mode=True, gamma_bar=1.0e-3, jac_exact=0):
    """Crystal properties of the FCC cubic Voce material (defaults: the
    C11/C12/C44 = 168.4/121.4/75.4 GPa material)."""
    if e is None:
        e = (C11 - C12) * (C11 + 2 * C12) / (C11 + C12)
        nu = C12 / (C11 + C12)
    c = OrcCrystal()
    rc = lib().orc_crystal_init(ctypes.byref(c), slip_set, e, nu, mu, rate_n, tau_y, tau_v,
                                theta_0, voche_m, int(alter_mode), gamma_bar)
    assert rc == 0
    c.jac_exact = int(jac_exact)
    return c


def voxel_update(c, angles, step, it, dt, Fn, Fn1, sn):
    ns = c.nslip
    snp1 = np.zeros(nwords(ns))
    Pv = np.zeros(9)
    K = np.zeros(81)
    st = np.zeros(5, dtype=np.int32)
    fail = lib().orc_voxel_update(ctypes.byref(c), P(np.ascontiguousarray(angles, float)),
                                  step, it, dt, P(np.ascontiguousarray(Fn.ravel(), float)),
                                  P(np.ascontiguousarray(Fn1.ravel(), float)),
                                  P(np.ascontiguousarray(sn, float)), P(snp1), P(Pv), P(K),
                                  IP(st))
    return Pv, K, snp1, st, fail


def state_init(c, V):
    s = np.zeros((nwords(c.nslip), V))
    lib().orc_state_init(ctypes.byref(c), V, P(s))
    return s


def drive_eps_sig(c, angles, step, it, dt, Fn, Fn1, sn):
    V = Fn.shape[1]
    snp1 = np.zeros_like(sn)
    Pn1 = np.zeros((9, V))
    K4 = np.zeros((81, V))
    st = np.zeros(5, dtype=np.int64)
    lib().orc_drive_eps_sig(ctypes.byref(c), V, P(np.ascontiguousarray(angles)), step, it, dt,
                            P(np.ascontiguousarray(Fn)), P(np.ascontiguousarray(Fn1)),
                            P(np.ascontiguousarray(sn)), P(snp1), P(Pn1), P(K4),
                            st.ctypes.data_as(_lp))
    return Pn1, K4, snp1, st


class FFT:
    def __init__(self, N, mode=0):
        self.N = N
        self.V = N ** 3
        self.h = lib().orc_fft_create2(N, mode)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_fft_destroy(self.h)
            self.h = None

    def gkdf(self, F, K4=None):
        out = np.zeros((9, self.V))
        if K4 is None:
            lib().orc_gkdf(self.h, None, P(np.ascontiguousarray(F)), P(out), 0)
        else:
            lib().orc_gkdf(self.h, P(np.ascontiguousarray(K4)), P(np.ascontiguousarray(F)),
                           P(out), 1)
        return out

    def pcg(self, K4, b, tol, maxit=None):
        """fftPcg; maxit bounds the iterations (the reference's 1000 if None)"""
        x = np.zeros((9, self.V))
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0)
        K4c, bc = np.ascontiguousarray(K4), np.ascontiguousarray(b)
        if maxit is None:
            rc = lib().orc_pcg(self.h, P(K4c), P(bc), P(x), tol, ctypes.byref(it), ctypes.byref(rel))
        else:
            rc = lib().orc_pcg_limit(self.h, P(K4c), P(bc), P(x), tol, int(maxit), ctypes.byref(it),
                                     ctypes.byref(rel))
        return x, rc, it.value, rel.value

    def tangent_homo(self, K4, tol):
        C = np.zeros(81)
        lib().orc_tangent_homo(self.h, P(np.ascontiguousarray(K4)), tol, P(C))
        return C


def nbc_update(C_homo, DbarF, P_bar, PBC, isNBC):
    d = np.ascontiguousarray(DbarF, float).copy()
    rc = lib().orc_nbc_update(P(np.ascontiguousarray(C_homo, float)), P(d),
                              P(np.ascontiguousarray(P_bar, float)),
                              P(np.ascontiguousarray(PBC, float)),
                              IP(np.ascontiguousarray(isNBC, np.int32)))
    return d, rc


def fft_nr3(c, N, angles, dt, BC_all, isNBC, tolPCG=1e-8, tolNR=1e-6, maxIter=25,
            gamma_mode=0):
    """Run the oracle FFT_nr3 loop; returns (Pbar[nstep,9], counters[nstep,4], rc, state)"""
    V = N ** 3
    nstep = BC_all.shape[0]
    eye = np.zeros((9, V))
    eye[[0, 4, 8]] = 1.0
    st = dict(Fn=eye.copy(), Fn1=eye.copy(), Pn=np.zeros((9, V)), Pn1=np.zeros((9, V)),
              K4=np.zeros((81, V)), sn=state_init(c, V), snp1=state_init(c, V))
    ang = np.ascontiguousarray(angles, float)
    pb = OrcProblem(N=N, V=V, cp=ctypes.pointer(c), angles=P(ang), dt=dt, tolPCG=tolPCG,
                    tolNR=tolNR, maxIter=maxIter, gamma_mode=gamma_mode, **{k: P(v) for k, v in st.items()})
    Pbar = np.zeros((nstep, 9))
    cnt = np.zeros((nstep, 4), dtype=np.int32)
    rc = lib().orc_fft_nr3(ctypes.byref(pb), nstep, P(np.ascontiguousarray(BC_all, float)),
                           IP(np.ascontiguousarray(isNBC, np.int32)), P(Pbar), IP(cnt))
    return Pbar, cnt, rc, st
