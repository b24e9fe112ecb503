"""cpfft_amd -- MI355X implementation of CPFFT's finite-strain FFT solver step.

Host-side Python mirror of the reference's operator interface over the
C-ABI of libcpfft_amd.so (include/cpfft.h) and the Fortran load-step driver
libcpfft_driver.so (cpfft_compute).  There is no CPU path: loading fails
loudly when the native libraries are missing, and the operators need a HIP
device.
"""
import ctypes
import os

import numpy as np

from .configs import WORKLOADS, Material, Workload  # noqa: F401

PKG = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(PKG, "libcpfft_amd.so")
DRIVER = os.path.join(PKG, "libcpfft_driver.so")

FN, FN1, PN, PN1, DFM, B, K4, STATE_N, STATE_NP1, WORK, EPS = range(11)
WARM_NETWORK, WARM_PREVIOUS = 1, 2  # include/cpfft.h CPFFT_WARM_*

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


class CMaterial(ctypes.Structure):
    _fields_ = [
        ("slip_set", ctypes.c_int32),
        ("e", ctypes.c_double), ("nu", ctypes.c_double), ("mu", ctypes.c_double),
        ("rate_n", ctypes.c_double), ("tau_y", ctypes.c_double), ("tau_v", ctypes.c_double),
        ("theta_0", ctypes.c_double), ("voche_m", ctypes.c_double),
        ("gamma_bar", ctypes.c_double), ("alter_mode", ctypes.c_int32),
        ("atol", ctypes.c_double), ("atol1", ctypes.c_double), ("rtol", ctypes.c_double),
        ("rtol1", ctypes.c_double), ("miter", ctypes.c_int32), ("jac_exact", ctypes.c_int32),
    ]


# every symbol include/cpfft.h declares
EXPORTS = [
    "cpfft_create", "cpfft_destroy", "cpfft_last_error", "cpfft_nslip", "cpfft_field",
    "cpfft_upload", "cpfft_download", "cpfft_download_voxels", "cpfft_reset", "cpfft_sync", "cpfft_driver_state",
    "cpfft_gkdf", "cpfft_pcg",
    "cpfft_drive_eps_sig", "cpfft_tangent_homo", "cpfft_update", "cpfft_dnrm2", "cpfft_daxpy",
    "cpfft_dscal", "cpfft_dcopy", "cpfft_fill9", "cpfft_mean9", "cpfft_f2disp",
    "cpfft_gkdf_noalloc", "cpfft_counters",
    "cpfft_kernel_times", "cpfft_set_timing", "cpfft_set_graphs", "cpfft_set_cg_maxit", "cpfft_autotune",
    "cpfft_get_tune", "cpfft_set_tune",
    "cpfft_mlp_init", "cpfft_mlp_log", "cpfft_mlp_train", "cpfft_mlp_warm_start",
    "cpfft_set_warm_start", "cpfft_local_stats",
    "cpfft_mlp_params", "cpfft_mlp_forward", "cpfft_mlp_grad", "cpfft_mlp_predict",
    "cpfft_mlp_stats",
]

# every symbol include/cpfft_driver.h declares (libcpfft_driver.so)
DRIVER_EXPORTS = ["cpfft_compute", "cpfft_compute_batch"]


class CBatchItem(ctypes.Structure):
    """cpfft_batch_item of include/cpfft_driver.h"""
    _fields_ = [
        ("ctx", _vp), ("first_step", ctypes.c_int), ("nstep", ctypes.c_int),
        ("BC_all", _dp), ("isNBC", _ip), ("tolPCG", ctypes.c_double),
        ("tolNR", ctypes.c_double), ("maxIter", ctypes.c_int), ("weight", ctypes.c_double),
        ("Pbar_out", _dp), ("counters", _ip), ("times", ctypes.c_double * 3),
        ("rc", ctypes.c_int),
    ]


MLP_IN, MLP_HID, MLP_OUT = 42, 128, 12
MLP_NPARAM, MLP_NNORM = 23564, 108


def mlp_sizes(nslip):
    """(nin, nout, nparam, nnorm) of the warm-start network for a crystal of
    nslip slip systems (include/cpfft.h CPFFT_MLP_NIN/NOUT/NPARAMS/NNORMS)"""
    nin = 18 + 2 * nslip
    h = MLP_HID
    return nin, nslip, h * (nin + 1) + h * (h + 1) + nslip * (h + 1), 2 * nin + 2 * nslip

_lib = None
_drv = None


def lib():
    """Load libcpfft_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build the HIP library first "
                               "(python -m cpfft_amd.build)")
        L = ctypes.CDLL(LIB)
        L.cpfft_create.argtypes = [ctypes.c_int, ctypes.POINTER(CMaterial), _dp, ctypes.c_double,
                                   ctypes.c_int, ctypes.POINTER(_vp)]
        L.cpfft_destroy.argtypes = [_vp]
        L.cpfft_last_error.restype = ctypes.c_char_p
        L.cpfft_nslip.argtypes = [_vp]
        L.cpfft_field.argtypes = [_vp, ctypes.c_int]
        L.cpfft_field.restype = _vp
        L.cpfft_set_graphs.argtypes = [_vp, ctypes.c_int]
        L.cpfft_set_cg_maxit.argtypes = [_vp, ctypes.c_int]
        L.cpfft_gkdf_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.cpfft_autotune.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.cpfft_get_tune.argtypes = [_vp, _ip]
        L.cpfft_set_tune.argtypes = [_vp, _ip]
        L.cpfft_driver_state.argtypes = [_vp]
        L.cpfft_driver_state.restype = _dp
        L.cpfft_upload.argtypes = [_vp, ctypes.c_int, _dp]
        L.cpfft_download.argtypes = [_vp, ctypes.c_int, _dp]
        L.cpfft_download_voxels.argtypes = [_vp, ctypes.c_int, ctypes.c_long, ctypes.c_long, _dp]
        L.cpfft_reset.argtypes = [_vp]
        L.cpfft_sync.argtypes = [_vp]
        L.cpfft_gkdf.argtypes = [_vp, _vp, _vp, ctypes.c_int]
        L.cpfft_pcg.argtypes = [_vp, _vp, _vp, ctypes.c_double, _ip, _dp]
        L.cpfft_drive_eps_sig.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.cpfft_tangent_homo.argtypes = [_vp, ctypes.c_double, _dp]
        L.cpfft_update.argtypes = [_vp]
        L.cpfft_dnrm2.argtypes = [_vp, _vp, _dp]
        L.cpfft_daxpy.argtypes = [_vp, ctypes.c_double, _vp, _vp]
        L.cpfft_dscal.argtypes = [_vp, ctypes.c_double, _vp]
        L.cpfft_dcopy.argtypes = [_vp, _vp, _vp]
        L.cpfft_fill9.argtypes = [_vp, _vp, _dp]
        L.cpfft_mean9.argtypes = [_vp, _vp, _dp]
        L.cpfft_f2disp.argtypes = [_vp, _vp, _dp, _dp]
        L.cpfft_counters.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong)]
        L.cpfft_kernel_times.argtypes = [_vp, _dp]
        L.cpfft_set_timing.argtypes = [_vp, ctypes.c_int]
        _fp = ctypes.POINTER(ctypes.c_float)
        L.cpfft_mlp_init.argtypes = [_vp, ctypes.c_uint64, ctypes.c_long, ctypes.c_int]
        L.cpfft_mlp_log.argtypes = [_vp, ctypes.c_int, ctypes.c_long]
        L.cpfft_mlp_train.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]
        L.cpfft_mlp_warm_start.argtypes = [_vp, ctypes.c_int]
        L.cpfft_set_warm_start.argtypes = [_vp, ctypes.c_int]
        L.cpfft_local_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong)]
        L.cpfft_mlp_params.argtypes = [_vp, _fp, ctypes.c_int]
        L.cpfft_mlp_forward.argtypes = [_vp, _fp, ctypes.c_int, _fp]
        L.cpfft_mlp_grad.argtypes = [_vp, _fp, _fp, ctypes.c_int, _fp, _dp]
        L.cpfft_mlp_predict.argtypes = [_vp, ctypes.c_int, _fp, _fp]
        L.cpfft_mlp_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong)]
        _lib = L
    return _lib


def driver():
    """Load the Fortran driver libcpfft_driver.so (raises if not built)."""
    global _drv
    if _drv is None:
        lib()
        if not os.path.exists(DRIVER):
            raise RuntimeError(f"{DRIVER} is missing: build it first (python -m cpfft_amd.build)")
        D = ctypes.CDLL(DRIVER)
        D.cpfft_compute.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _ip, _dp,
                                    ctypes.c_int, ctypes.c_int]
        D.cpfft_compute_batch.argtypes = [ctypes.POINTER(CBatchItem), ctypes.c_int, ctypes.c_int]
        _drv = D
    return _drv


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _pf(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class CpfftError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        raise CpfftError(f"{what} failed ({rc}): {lib().cpfft_last_error().decode()}")
    return rc


class Solver:
    """One RVE on one GPU: the reference's module `fft` state in HBM plus its
    operators (G_K_dF, fftPcg, drive_eps_sig, tangent_homo, update)."""

    def __init__(self, N, angles, material=None, dt=0.5, gamma_mode=1):
        m = material or Material()
        cm = CMaterial(**{k: getattr(m, k) for k, _ in CMaterial._fields_})
        self.N = N
        self.V = N ** 3
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        assert ang.shape == (3, self.V), "angles must be [3][N^3]"
        h = _vp()
        _check(lib().cpfft_create(N, ctypes.byref(cm), _p(ang), dt, gamma_mode, ctypes.byref(h)),
               "cpfft_create")
        self.h = h
        self.nslip = lib().cpfft_nslip(h)
        self.nwords = 26 + 2 * self.nslip

    def close(self):
        if getattr(self, "h", None):
            lib().cpfft_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def field(self, f):
        return lib().cpfft_field(self.h, f)

    def _len(self, f):
        return {K4: 81, STATE_N: self.nwords, STATE_NP1: self.nwords, EPS: 6}.get(f, 9) * self.V

    def upload(self, f, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self._len(f)
        _check(lib().cpfft_upload(self.h, f, _p(a)), "cpfft_upload")

    def download(self, f):
        out = np.empty(self._len(f))
        _check(lib().cpfft_download(self.h, f, _p(out)), "cpfft_download")
        rows = self._len(f) // self.V
        return out.reshape(rows, self.V)

    def download_voxels(self, f, e0, n):
        """voxels [e0, e0+n) of every component of field f, [rows][n]"""
        rows = self._len(f) // self.V
        out = np.empty(rows * n)
        _check(lib().cpfft_download_voxels(self.h, f, int(e0), int(n), _p(out)),
               "cpfft_download_voxels")
        return out.reshape(rows, n)

    def reset(self):
        _check(lib().cpfft_reset(self.h), "cpfft_reset")

    def gkdf(self, src, dst, flgK):
        _check(lib().cpfft_gkdf(self.h, self.field(src), self.field(dst), int(flgK)), "cpfft_gkdf")

    def pcg(self, b, x, tol):
        it = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        rc = lib().cpfft_pcg(self.h, self.field(b), self.field(x), tol, ctypes.byref(it),
                             ctypes.byref(rel))
        if rc < -1:
            raise CpfftError(lib().cpfft_last_error().decode())
        return rc, it.value, rel.value

    def drive_eps_sig(self, step, iiter):
        return _check(lib().cpfft_drive_eps_sig(self.h, step, iiter), "cpfft_drive_eps_sig")

    def tangent_homo(self, tol):
        C = np.zeros(81)
        _check(lib().cpfft_tangent_homo(self.h, tol, _p(C)), "cpfft_tangent_homo")
        return C

    def update(self):
        _check(lib().cpfft_update(self.h), "cpfft_update")

    # BLAS-1 on whole [9][V] fields (the MKL calls of FFT_nr3.f:86-147)
    def dnrm2(self, f):
        out = ctypes.c_double(0.0)
        _check(lib().cpfft_dnrm2(self.h, self.field(f), ctypes.byref(out)), "cpfft_dnrm2")
        return out.value

    def daxpy(self, a, x, y):
        _check(lib().cpfft_daxpy(self.h, a, self.field(x), self.field(y)), "cpfft_daxpy")

    def dscal(self, a, x):
        _check(lib().cpfft_dscal(self.h, a, self.field(x)), "cpfft_dscal")

    def dcopy(self, x, y):
        _check(lib().cpfft_dcopy(self.h, self.field(x), self.field(y)), "cpfft_dcopy")

    def fill9(self, x, v):
        vv = np.ascontiguousarray(v, dtype=np.float64)
        assert vv.size == 9
        _check(lib().cpfft_fill9(self.h, self.field(x), _p(vv)), "cpfft_fill9")

    def mean9(self, x):
        out = np.zeros(9)
        _check(lib().cpfft_mean9(self.h, self.field(x), _p(out)), "cpfft_mean9")
        return out

    def f2disp(self, f=None, lengths=(1.0, 1.0, 1.0)):
        """Nodal displacements [(N+1)^3][3] of the brick mesh from the
        deformation gradient in field f (default FN1), as f2disp.f."""
        P = self.N + 1
        u = np.empty(3 * P ** 3)
        ln = np.ascontiguousarray(lengths, dtype=np.float64)
        assert ln.size == 3
        _check(lib().cpfft_f2disp(self.h, self.field(FN1 if f is None else f), _p(ln), _p(u)),
               "cpfft_f2disp")
        return u.reshape(P ** 3, 3)

    def sync(self):
        _check(lib().cpfft_sync(self.h), "cpfft_sync")

    def counters(self):
        out = (ctypes.c_longlong * 8)()
        _check(lib().cpfft_counters(self.h, out), "cpfft_counters")
        return list(out)

    def set_graphs(self, on):
        """replay CG iteration chunks from HIP graphs (results unchanged):
        True or 1 captures the driver's graphs now, 2 at first use"""
        _check(lib().cpfft_set_graphs(self.h, int(on)), "cpfft_set_graphs")

    def set_cg_maxit(self, maxit):
        """iteration limit of the CG solves (fftPcg's maxIter, 1000)"""
        _check(lib().cpfft_set_cg_maxit(self.h, int(maxit)), "cpfft_set_cg_maxit")

    def autotune(self, reps=10, force=False):
        """choose the fastest launch shapes of the planned FFT path (results
        unchanged); returns the shapes in use (see tune())"""
        rc = lib().cpfft_autotune(self.h, int(reps), int(force))
        if rc < 0:
            raise CpfftError(lib().cpfft_last_error().decode())
        return self.tune()

    def tune(self):
        """[pencils per block of pass A with K4, without K4, lines per block
        of the y passes, y lines per block of pass C]"""
        out = (ctypes.c_int * 4)()
        _check(lib().cpfft_get_tune(self.h, out), "cpfft_get_tune")
        return list(out)

    def set_tune(self, shapes):
        arr = (ctypes.c_int * 4)(*[int(v) for v in shapes])
        _check(lib().cpfft_set_tune(self.h, arr), "cpfft_set_tune")

    def set_timing(self, on):
        """0/False off; True or 1 every launch group; k > 1 every k-th."""
        _check(lib().cpfft_set_timing(self.h, int(on)), "cpfft_set_timing")

    def kernel_times(self):
        """(average ms per launch group, counts) for: G_K_dF with K4,
        drive_eps_sig, CG vector update, G_K_dF without K4"""
        out = np.zeros(8)
        lib().cpfft_kernel_times(self.h, _p(out))
        return out[:4], out[4:].astype(int)

    # ---- learned warm start (no counterpart in the reference) ----------
    def mlp_init(self, seed=1, capacity=1 << 20, batch=4096):
        _check(lib().cpfft_mlp_init(self.h, seed, capacity, batch), "cpfft_mlp_init")

    def mlp_train(self, iters, lr=1e-3, seed=7):
        loss = ctypes.c_double(0.0)
        _check(lib().cpfft_mlp_train(self.h, iters, lr, seed, ctypes.byref(loss)), "cpfft_mlp_train")
        return loss.value

    def mlp_warm_start(self, on):
        _check(lib().cpfft_mlp_warm_start(self.h, int(on)), "cpfft_mlp_warm_start")

    def set_warm_start(self, mode):
        """Warm-start sources of the local solves: WARM_NETWORK | WARM_PREVIOUS
        (0: the reference's start)."""
        _check(lib().cpfft_set_warm_start(self.h, int(mode)), "cpfft_set_warm_start")

    def mlp_params(self, new=None):
        """Get (new=None) or set the flat parameters + normalisation vector."""
        if new is None:
            _, _, npar, nnorm = mlp_sizes(self.nslip)
            out = np.zeros(npar + nnorm, dtype=np.float32)
            _check(lib().cpfft_mlp_params(self.h, _pf(out), 1), "cpfft_mlp_params")
            return out
        buf = np.ascontiguousarray(new, dtype=np.float32)
        _, _, npar, nnorm = mlp_sizes(self.nslip)
        assert buf.size == npar + nnorm
        _check(lib().cpfft_mlp_params(self.h, _pf(buf), 0), "cpfft_mlp_params")

    def mlp_forward(self, Xn):
        X = np.ascontiguousarray(Xn, dtype=np.float32)
        Y = np.zeros((X.shape[0], mlp_sizes(self.nslip)[1]), dtype=np.float32)
        _check(lib().cpfft_mlp_forward(self.h, _pf(X), X.shape[0], _pf(Y)), "cpfft_mlp_forward")
        return Y

    def mlp_grad(self, Xn, Tn):
        X = np.ascontiguousarray(Xn, dtype=np.float32)
        T = np.ascontiguousarray(Tn, dtype=np.float32)
        g = np.zeros(mlp_sizes(self.nslip)[2], dtype=np.float32)
        loss = ctypes.c_double(0.0)
        _check(lib().cpfft_mlp_grad(self.h, _pf(X), _pf(T), X.shape[0], _pf(g), ctypes.byref(loss)),
               "cpfft_mlp_grad")
        return g, loss.value

    def mlp_predict(self, step):
        """(raw network inputs [V][nin], predicted slip increments as
        normalised resolved shears [nout][V]) for the current state"""
        nin, nout, _, _ = mlp_sizes(self.nslip)
        X = np.zeros((self.V, nin), dtype=np.float32)
        ws = np.zeros((nout, self.V), dtype=np.float32)
        _check(lib().cpfft_mlp_predict(self.h, int(step), _pf(X), _pf(ws)), "cpfft_mlp_predict")
        return X, ws

    def local_stats(self):
        """(voxel local iterations, issued local iterations, line-search
        evaluations, sub-step cuts) since creation"""
        out = (ctypes.c_longlong * 4)()
        _check(lib().cpfft_local_stats(self.h, out), "cpfft_local_stats")
        return list(out)

    def mlp_stats(self):
        out = (ctypes.c_longlong * 4)()
        _check(lib().cpfft_mlp_stats(self.h, out), "cpfft_mlp_stats")
        return list(out)

    def compute(self, BC_all, isNBC, tolPCG=1e-8, tolNR=1e-6, maxIter=25, verbose=False,
                first_step=1, mlp_log_last=0, mlp_log_stride=1):
        """The `compute` command: drive_eps_sig(1,0) then the FFT_nr3 load-step
        loop, run by the Fortran driver, for the steps whose boundary values
        BC_all holds, numbered from first_step (a first_step above 1
        continues the previous call).  Steps up to mlp_log_last log
        warm-start training pairs (every mlp_log_stride-th voxel).  Returns
        (rc, Pbar[nstep,9], counters[nstep,4], times[3])."""
        BC = np.ascontiguousarray(BC_all, dtype=np.float64)
        nstep = BC.shape[0]
        nbc = np.ascontiguousarray(isNBC, dtype=np.int32)
        Pbar = np.zeros((nstep, 9))
        cnt = np.zeros((nstep, 4), dtype=np.int32)
        times = np.zeros(3)
        rc = driver().cpfft_compute(self.h, first_step, nstep, _p(BC), nbc.ctypes.data_as(_ip),
                                    tolPCG, tolNR,
                                    maxIter, int(verbose), _p(Pbar), cnt.ctypes.data_as(_ip),
                                    _p(times), int(mlp_log_last), int(mlp_log_stride))
        return rc, Pbar, cnt, times

    def driver_state(self):
        """(barF_t[9], P_bar[9], C_homo[81]) the driver carries between steps"""
        p = lib().cpfft_driver_state(self.h)
        a = np.ctypeslib.as_array(p, shape=(99,)).copy()
        return a[0:9], a[9:18], a[18:99]


def compute_batch(jobs, nthreads=8):
    """Batched uneven RVEs (cpfft_compute_batch): jobs is a list of
    (solver, BC_all, isNBC, tolPCG, tolNR, maxIter, first_step).  Runs
    them on nthreads host threads, heaviest first, each RVE on its own
    context and stream.  Returns (rc, [(rc, Pbar, counters, times), ...])
    in the order of jobs."""
    n = len(jobs)
    items = (CBatchItem * max(n, 1))()
    keep = []
    out = []
    for k, (s, BC_all, isNBC, tolPCG, tolNR, maxIter, first_step) in enumerate(jobs):
        BC = np.ascontiguousarray(BC_all, dtype=np.float64)
        nbc = np.ascontiguousarray(isNBC, dtype=np.int32)
        nstep = BC.shape[0]
        Pbar = np.zeros((nstep, 9))
        cnt = np.zeros((nstep, 4), dtype=np.int32)
        keep += [BC, nbc]
        out.append((Pbar, cnt))
        it = items[k]
        it.ctx, it.first_step, it.nstep = s.h, first_step, nstep
        it.BC_all, it.isNBC = _p(BC), nbc.ctypes.data_as(_ip)
        it.tolPCG, it.tolNR, it.maxIter = tolPCG, tolNR, maxIter
        it.weight = float(s.N) ** 3 * nstep
        it.Pbar_out, it.counters = _p(Pbar), cnt.ctypes.data_as(_ip)
    rc = driver().cpfft_compute_batch(items, n, int(nthreads))
    res = [(items[k].rc, out[k][0], out[k][1], np.array(items[k].times[:])) for k in range(n)]
    return rc, res
