"""cunfft_amd -- MI355X-native NFFT behind the CUNFFT plan interface.

Python mirror of the reference's plan surface (src/cunfft.h:35-261) over the
C ABI of ``libcunfft_amd.so`` (``include/cunfft.h``).  The call sequence is
the reference's (``cunfftDoc.pdf`` p.9, ``src/simpleTest.cpp:18-80``)::

    p = Plan([N0, N1], M)              # cunfft_init / cunfft_initGuru_m
    p.x[:] = ...; p.f_hat[:] = ...     # fill the pinned host arrays
    p.copyDataToDevice()
    p.transform()                      # cunfft_transform
    p.copyDataToHost()                 # result in p.f

Errors become :class:`CunfftError` (the library records a status per plan and
never exits).  There is no CPU path: if the shared library is missing the
import of :func:`lib` fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = ["Plan", "CunfftError", "lib", "LIB_PATH", "CUFFT_INIT", "CUFFT_BOTH",
           "CUFFT_FORW", "CUFFT_BACKW", "MEASURE_TIME", "NO_HOST_BUFFERS",
           "ATOMIC_SCATTER", "PLAIN_GATHER", "NODE_LANES", "default_n"]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcunfft_amd.so")

CUFFT_INIT = 1 << 0
CUFFT_BOTH = 1 << 1
CUFFT_FORW = 1 << 2
CUFFT_BACKW = 1 << 3
MEASURE_TIME = 1 << 8
NO_HOST_BUFFERS = 1 << 9
ATOMIC_SCATTER = 1 << 10
PLAIN_GATHER = 1 << 11
NODE_LANES = 1 << 12
COMPENSATED = 1 << 13
ASYNC_COPY = 1 << 14
KAISER_BESSEL = 1 << 15
SINC = 1 << 16
DEFAULT_CUT_OFF = 6
TILE_SHAPES = 3                     # plane-tiled configurations 0 .. TILE_SHAPES-1
TUNE_CONFIGS = TILE_SHAPES + 1      # + the node-lane configuration
TUNE_FORWARD = 1
TUNE_ADJOINT = 2

_STATUS = {0: "success", 1: "invalid argument", 2: "unsupported parameters",
           3: "allocation failed", 4: "HIP runtime error", 5: "invalid plan state"}


class CunfftError(RuntimeError):
    def __init__(self, status: int, message: str = ""):
        self.status = status
        super().__init__(f"cunfft status {status} ({_STATUS.get(status, '?')}): {message}")


class NFFTTimeSpec(ctypes.Structure):
    """src/timeMeasurement.h:22-36."""
    _fields_ = [(k, ctypes.c_double) for k in
                ("runTime", "time_ROC", "time_CONV", "time_FFT", "time_COPY_IN",
                 "time_COPY_OUT", "time_else")]


def _plan_fields():
    P = ctypes.POINTER
    return [
        ("N_total", ctypes.c_ssize_t), ("M_total", ctypes.c_ssize_t), ("d", ctypes.c_int),
        ("N", P(ctypes.c_ssize_t)), ("x", P(ctypes.c_double)), ("f_hat", ctypes.c_void_p),
        ("f", ctypes.c_void_p), ("sigma", P(ctypes.c_double)), ("n", P(ctypes.c_ssize_t)),
        ("n_total", ctypes.c_ssize_t), ("m", ctypes.c_int), ("b", P(ctypes.c_double)),
        ("g", ctypes.c_void_p), ("g_hat", ctypes.c_void_p), ("fg_exp_l_cpu", P(ctypes.c_double)),
        ("g_gpu", ctypes.c_void_p), ("fhat_gpu", ctypes.c_void_p), ("x_gpu", ctypes.c_void_p),
        ("f_gpu", ctypes.c_void_p), ("fft_forward", ctypes.c_void_p),
        ("fft_backward", ctypes.c_void_p), ("CUNFFTTimes", NFFTTimeSpec),
        ("flags", ctypes.c_uint), ("impl", ctypes.c_void_p), ("status", ctypes.c_int),
    ]


class cunfft_plan(ctypes.Structure):
    """Layout of ``cunfft_plan`` / ``cunfftf_plan`` (src/cunfft_typedefs.h:118-178 + impl, status)."""
    _fields_ = _plan_fields()


_lib = None

# (name suffix, FP64 symbol, FP32 symbol)
_PLAN_VOID = {
    "copyDataToDevice": ("copyDataToDevice", "cunfftf_copyDataToDevice"),
    "copyDataToDeviceAd": ("copyDataToDeviceAd", "cunfftf_copyDataToDeviceAd"),
    "copyDataToHost": ("copyDataToHost", "cunfftf_copyDataToHost"),
    "copyDataToHostAd": ("copyDataToHostAd", "cunfftf_copyDataToHostAd"),
    "copy_g_ToHost": ("copy_g_ToHost", "cunfftf_copy_g_ToHost"),
    "transform": ("cunfft_transform", "cunfftf_transform"),
    "adjoint": ("cunfft_adjoint", "cunfftf_adjoint"),
    "cundft_transform": ("cundft_transform", "cunfftf_ndft_transform"),
    "cundft_adjoint": ("cundft_adjoint", "cunfftf_ndft_adjoint"),
    "reinit": ("cunfft_reinit", "cunfftf_reinit"),
    "reinitAd": ("cunfft_reinitAd", "cunfftf_reinitAd"),
    "finalizeDevice": ("cunfft_finalizeDevice", "cunfftf_finalizeDevice"),
    "initDevicePara": ("cunfft_initDevicePara", "cunfftf_initDevicePara"),
    "freePlan": ("cunfft_freePlan", "cunfftf_freePlan"),
    "copy_f_ToHost": ("copy_f_ToHost", "cunfftf_copy_f_ToHost"),
    "copy_f_hat_ToHost": ("copy_f_hat_ToHost", "cunfftf_copy_f_hat_ToHost"),
    "copy_x_toHost": ("copy_x_toHost", "cunfftf_copy_x_toHost"),
}
_PLAN_PTR_ARG = ("copy_f_ToHost", "copy_f_hat_ToHost", "copy_x_toHost")


def lib():
    """Load libcunfft_amd.so (raises if it was not built -- there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing; build it with `python -m cunfft_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    PP = ctypes.POINTER(cunfft_plan)
    vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_ssize_t)
    for k, (a, b) in _PLAN_VOID.items():
        for s in (a, b):
            getattr(L, s).argtypes = [PP, vp] if k in _PLAN_PTR_ARG else [PP]
            getattr(L, s).restype = None
    for s in ("cunfft_finalize", "cunfftf_finalize"):
        getattr(L, s).argtypes = [PP]
        getattr(L, s).restype = None
    for s in ("cunfft_initGuru_m", "cunfftf_initGuru_m"):
        getattr(L, s).argtypes = [PP, ctypes.c_int, i64p, ctypes.c_ssize_t, i64p, ctypes.c_int,
                                  ctypes.c_uint]
        getattr(L, s).restype = None
    for s in ("cunfft_last_error", "cunfftf_last_error"):
        getattr(L, s).argtypes = [PP]
        getattr(L, s).restype = ctypes.c_int
    for s in ("copy_f_hat_ToHost", "copy_f_ToHost", "cunfftf_copy_f_hat_ToHost",
              "cunfftf_copy_f_ToHost", "copy_x_toHost", "cunfftf_copy_x_toHost"):
        getattr(L, s).argtypes = [PP, vp]
        getattr(L, s).restype = None
    for s in ("cunfft_set_nodes_dev", "cunfftf_set_nodes_dev"):
        getattr(L, s).argtypes = [PP, vp, vp]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_transform_dev", "cunfftf_transform_dev", "cunfft_adjoint_dev",
              "cunfftf_adjoint_dev"):
        getattr(L, s).argtypes = [PP, vp, vp, vp]
        getattr(L, s).restype = ctypes.c_int
    L.cunfft_adjoint_noalloc_dev.argtypes = [PP, vp, vp, vp]
    L.cunfft_adjoint_noalloc_dev.restype = ctypes.c_int
    for s in ("cunfft_grad_x_dev", "cunfftf_grad_x_dev"):
        getattr(L, s).argtypes = [PP, vp, vp, vp, vp]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_backward_dev", "cunfftf_backward_dev"):
        getattr(L, s).argtypes = [PP, vp, vp, vp, vp, vp]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cundft_transform_rows_dev", "cundft_adjoint_cols_dev",
              "cunfftf_ndft_transform_rows_dev", "cunfftf_ndft_adjoint_cols_dev"):
        getattr(L, s).argtypes = [PP, vp, vp, ctypes.c_longlong, vp, vp]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_solve_dev", "cunfftf_solve_dev"):
        getattr(L, s).argtypes = [PP, vp, vp, vp, ctypes.c_int, ctypes.c_double,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), vp]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_synchronize", "cunfftf_synchronize"):
        getattr(L, s).argtypes = [PP]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_autotune", "cunfftf_autotune"):
        getattr(L, s).argtypes = [PP, ctypes.c_int, ctypes.c_int, vp, vp]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_get_config", "cunfftf_get_config"):
        getattr(L, s).argtypes = [PP, ctypes.c_int]
        getattr(L, s).restype = ctypes.c_int
    for s in ("cunfft_set_config", "cunfftf_set_config"):
        getattr(L, s).argtypes = [PP, ctypes.c_int, ctypes.c_int]
        getattr(L, s).restype = ctypes.c_int
    L.cunfft_last_message.restype = ctypes.c_char_p
    L.cunfft_error_string.restype = ctypes.c_char_p
    L.cunfft_error_string.argtypes = [ctypes.c_int]
    _lib = L
    return L


def default_n(N):
    """``cunfft_init``'s rule n_t = 2 * next_pow2(N_t) (src/cunfft.cu:296-299)."""
    out = []
    for v in np.atleast_1d(N):
        p = 1
        while p < int(v):
            p <<= 1
        out.append(2 * p)
    return out


class Plan:
    """One NFFT plan (``cunfft_plan`` for FP64, ``cunfftf_plan`` for FP32).

    Parameters follow ``cunfft_initGuru`` (src/cunfft.h:168) plus the runtime
    cut-off ``m``: ``N`` bandwidths (length d), ``M`` nodes, ``n`` oversampled
    grid (default ``2*next_pow2(N)``), ``m`` window cut-off, ``precision``
    ``"double"`` or ``"single"``.
    """

    def __init__(self, N, M, n=None, m=DEFAULT_CUT_OFF, precision="double", flags=None):
        L = lib()
        N = [int(v) for v in np.atleast_1d(N)]
        self.d = len(N)
        self.precision = {"double": "double", "f64": "double", "single": "single",
                          "f32": "single", "float": "single"}[precision]
        self._sp = self.precision == "single"
        self._ctype = np.complex64 if self._sp else np.complex128
        if flags is None:
            flags = CUFFT_INIT | CUFFT_BOTH
        self._p = cunfft_plan()
        Narr = (ctypes.c_ssize_t * max(1, self.d))(*N)
        nptr = None
        if n is not None:
            narr = (ctypes.c_ssize_t * self.d)(*[int(v) for v in np.atleast_1d(n)])
            nptr = narr
        fn = L.cunfftf_initGuru_m if self._sp else L.cunfft_initGuru_m
        fn(ctypes.byref(self._p), self.d, Narr, int(M), nptr, int(m), int(flags))
        self._alive = True
        self._check()

    # ---------------------------------------------------------------- state
    def _check(self):
        st = self._p.status
        if st != 0:
            msg = lib().cunfft_last_message().decode(errors="replace")
            raise CunfftError(st, msg)

    def _call(self, name, *args):
        a, b = _PLAN_VOID[name]
        getattr(lib(), b if self._sp else a)(ctypes.byref(self._p), *args)
        self._check()

    @property
    def status(self):
        return self._p.status

    @property
    def N(self):
        return [self._p.N[t] for t in range(self.d)]

    @property
    def n(self):
        return [self._p.n[t] for t in range(self.d)]

    @property
    def m(self):
        return self._p.m

    @property
    def M_total(self):
        return self._p.M_total

    @property
    def N_total(self):
        return self._p.N_total

    @property
    def n_total(self):
        return self._p.n_total

    @property
    def sigma(self):
        return [self._p.sigma[t] for t in range(self.d)]

    @property
    def b(self):
        return [self._p.b[t] for t in range(self.d)]

    @property
    def flags(self):
        return self._p.flags

    @property
    def times(self):
        t = self._p.CUNFFTTimes
        return {k: getattr(t, k) for k, _ in NFFTTimeSpec._fields_}

    def _host(self, name, count, dtype):
        ptr = getattr(self._p, name)
        addr = ptr if isinstance(ptr, int) or ptr is None else ctypes.cast(ptr, ctypes.c_void_p).value
        if not addr:
            raise CunfftError(5, f"plan has no host array {name}")
        buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(addr)
        return np.frombuffer(buf, dtype=dtype, count=count)

    @property
    def x(self):
        """Pinned host nodes, shape (M, d), the reference's x[d*j+t]."""
        return self._host("x", self.M_total * self.d, np.float64).reshape(self.M_total, self.d)

    @property
    def f_hat(self):
        return self._host("f_hat", self.N_total, self._ctype)

    @property
    def f(self):
        return self._host("f", self.M_total, self._ctype)

    @property
    def g(self):
        return self._host("g", self.n_total, self._ctype)

    @property
    def device_pointers(self):
        p = self._p
        return {"g": p.g_gpu, "f_hat": p.fhat_gpu, "x": p.x_gpu, "f": p.f_gpu}

    # ------------------------------------------------- reference entry points
    def copyDataToDevice(self):
        self._call("copyDataToDevice")

    def copyDataToDeviceAd(self):
        self._call("copyDataToDeviceAd")

    def copyDataToHost(self):
        self._call("copyDataToHost")

    def copyDataToHostAd(self):
        self._call("copyDataToHostAd")

    def copy_g_ToHost(self):
        self._call("copy_g_ToHost")

    def copy_f_ToHost(self):
        out = np.empty(self.M_total, self._ctype)
        self._call("copy_f_ToHost", out.ctypes.data_as(ctypes.c_void_p))
        return out

    def copy_f_hat_ToHost(self):
        out = np.empty(self.N_total, self._ctype)
        self._call("copy_f_hat_ToHost", out.ctypes.data_as(ctypes.c_void_p))
        return out

    def copy_x_toHost(self):
        out = np.empty((self.M_total, self.d), np.float64)
        self._call("copy_x_toHost", out.ctypes.data_as(ctypes.c_void_p))
        return out

    def transform(self):
        self._call("transform")

    def adjoint(self):
        self._call("adjoint")

    def cundft_transform(self):
        self._call("cundft_transform")

    def cundft_adjoint(self):
        self._call("cundft_adjoint")

    def reinit(self):
        self._call("reinit")

    def reinitAd(self):
        self._call("reinitAd")

    def finalizeDevice(self):
        self._call("finalizeDevice")

    def initDevicePara(self):
        self._call("initDevicePara")

    def finalize(self):
        if getattr(self, "_alive", False):
            fn = lib().cunfftf_finalize if self._sp else lib().cunfft_finalize
            fn(ctypes.byref(self._p))
            self._alive = False

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.finalize()

    # ------------------------------------------------ device-pointer entries
    def _dev(self, a, b, *args):
        st = getattr(lib(), b if self._sp else a)(ctypes.byref(self._p), *args)
        if st != 0:
            raise CunfftError(st, lib().cunfft_last_message().decode(errors="replace"))

    def set_nodes_dev(self, x_ptr, stream=None):
        self._dev("cunfft_set_nodes_dev", "cunfftf_set_nodes_dev", ctypes.c_void_p(x_ptr),
                  ctypes.c_void_p(stream))

    def transform_dev(self, fhat_ptr, f_ptr, stream=None):
        self._dev("cunfft_transform_dev", "cunfftf_transform_dev", ctypes.c_void_p(fhat_ptr),
                  ctypes.c_void_p(f_ptr), ctypes.c_void_p(stream))

    def adjoint_dev(self, f_ptr, fhat_ptr, stream=None):
        self._dev("cunfft_adjoint_dev", "cunfftf_adjoint_dev", ctypes.c_void_p(f_ptr),
                  ctypes.c_void_p(fhat_ptr), ctypes.c_void_p(stream))

    def adjoint_noalloc_dev(self, f_ptr, fhat_ptr, stream=None):
        """The adjoint with no device allocation (cunfft_adjoint_noalloc_dev;
        FP64 plans): its scratch is the library's per-GPU array."""
        if self._sp:
            raise CunfftError(1, "cunfft_adjoint_noalloc_dev is an FP64 entry point")
        self._dev("cunfft_adjoint_noalloc_dev", "cunfft_adjoint_noalloc_dev", ctypes.c_void_p(f_ptr),
                  ctypes.c_void_p(fhat_ptr), ctypes.c_void_p(stream))

    def grad_x_dev(self, fhat_ptr, v_ptr, gx_ptr, stream=None):
        self._dev("cunfft_grad_x_dev", "cunfftf_grad_x_dev", ctypes.c_void_p(fhat_ptr),
                  ctypes.c_void_p(v_ptr), ctypes.c_void_p(gx_ptr), ctypes.c_void_p(stream))

    def backward_dev(self, fhat_ptr, grad_f_ptr, grad_fhat_ptr, grad_x_ptr, stream=None):
        """grad_fhat = A^H grad_f and grad_x = Re(conj(grad_f) dA/dx fhat); a 0 pointer
        skips that output (cunfft_backward_dev)."""
        self._dev("cunfft_backward_dev", "cunfftf_backward_dev", ctypes.c_void_p(fhat_ptr),
                  ctypes.c_void_p(grad_f_ptr), ctypes.c_void_p(grad_fhat_ptr),
                  ctypes.c_void_p(grad_x_ptr), ctypes.c_void_p(stream))

    def solve_dev(self, y_ptr, fhat_ptr, w_ptr=0, max_iter=50, tol=1e-10, stream=None):
        """Inverse NFFT by CGNR (cunfft_solve_dev): fhat (initial guess, then the
        solution) minimising sum_j w_j |y_j - (A fhat)_j|^2.  Returns (iterations,
        relative residual per iteration)."""
        hist = (ctypes.c_double * (max_iter + 1))()
        it = ctypes.c_int(0)
        self._dev("cunfft_solve_dev", "cunfftf_solve_dev", ctypes.c_void_p(y_ptr),
                  ctypes.c_void_p(w_ptr), ctypes.c_void_p(fhat_ptr), ctypes.c_int(max_iter),
                  ctypes.c_double(tol), hist, ctypes.byref(it), ctypes.c_void_p(stream))
        return it.value, [hist[k] for k in range(it.value + 1)]

    def synchronize(self):
        """Wait for the plan's queued work (ASYNC_COPY plans; cunfft_synchronize)."""
        self._dev("cunfft_synchronize", "cunfftf_synchronize")

    # ------------------------------------------------------------ autotuner
    def autotune(self, directions=TUNE_FORWARD | TUNE_ADJOINT, reps=5, stream=None):
        """Time the gridding configurations on the plan's current nodes and keep
        the fastest per direction, or, when both directions are tuned to
        different ones, the fastest forward + adjoint step of that pair and of
        either one on both directions, the pair only if its step is more than 2%
        faster (cunfft_autotune).  Returns ((forward config,
        adjoint config), forward times, adjoint times) with times in ms per
        configuration, -1 where a configuration does not apply or a direction
        was not tuned."""
        ms = (ctypes.c_float * (2 * TUNE_CONFIGS))()
        self._dev("cunfft_autotune", "cunfftf_autotune", ctypes.c_int(directions),
                  ctypes.c_int(reps), ctypes.cast(ms, ctypes.c_void_p), ctypes.c_void_p(stream))
        v = [float(t) for t in ms]
        return self.config, v[:TUNE_CONFIGS], v[TUNE_CONFIGS:]

    @property
    def config(self):
        """(forward, adjoint) gridding kernel configurations (cunfft_get_config)."""
        fn = lib().cunfftf_get_config if self._sp else lib().cunfft_get_config
        return (int(fn(ctypes.byref(self._p), TUNE_FORWARD)),
                int(fn(ctypes.byref(self._p), TUNE_ADJOINT)))

    @config.setter
    def config(self, c):
        """An int sets both directions; (forward, adjoint) sets each."""
        pairs = [(TUNE_FORWARD | TUNE_ADJOINT, c)] if isinstance(c, int) else \
            [(TUNE_FORWARD, c[0]), (TUNE_ADJOINT, c[1])]
        for dirs, v in pairs:
            self._dev("cunfft_set_config", "cunfftf_set_config", ctypes.c_int(dirs), ctypes.c_int(v))

    def ndft_rows_dev(self, fhat_ptr, rows_ptr, n_rows, out_ptr, stream=None):
        self._dev("cundft_transform_rows_dev", "cunfftf_ndft_transform_rows_dev",
                  ctypes.c_void_p(fhat_ptr), ctypes.c_void_p(rows_ptr),
                  ctypes.c_longlong(n_rows), ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream))

    def ndft_cols_dev(self, f_ptr, cols_ptr, n_cols, out_ptr, stream=None):
        self._dev("cundft_adjoint_cols_dev", "cunfftf_ndft_adjoint_cols_dev",
                  ctypes.c_void_p(f_ptr), ctypes.c_void_p(cols_ptr),
                  ctypes.c_longlong(n_cols), ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream))
