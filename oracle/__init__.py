"""CPU oracle for the CUNFFT hot path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``oracle/liboracle.so`` (an FP64 C restatement of the
reference algorithm, see ``oracle.h`` for the file:line map) with numpy.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product package
``cunfft_amd`` never imports it and has no CPU fallback.

Complex arrays are numpy ``complex128``; node arrays are ``float64`` of shape
``(M, d)`` (the reference's interleaved ``x[d*j+t]``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_drand48.restype = ctypes.c_double
        L.orc_drand48.argtypes = [_u64p]
        L.orc_shape_b.restype = ctypes.c_double
        L.orc_shape_b.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        L.orc_window_start.restype = ctypes.c_int64
        L.orc_window_start.argtypes = [ctypes.c_int64, ctypes.c_double, ctypes.c_int]
        L.orc_phi.restype = ctypes.c_double
        L.orc_phi.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_set_window.argtypes = [ctypes.c_int]
        L.orc_get_window.restype = ctypes.c_int
        L.orc_kb_shape.restype = ctypes.c_double
        L.orc_kb_shape.argtypes = [ctypes.c_int64, ctypes.c_int64]
        for fn in (L.orc_phi_kb, L.orc_dphi_kb):
            fn.restype = ctypes.c_double
            fn.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_bessel_i0.restype = ctypes.c_double
        L.orc_bessel_i0.argtypes = [ctypes.c_double]
        L.orc_sinc_shape.restype = ctypes.c_double
        L.orc_sinc_shape.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
        for fn in (L.orc_phi_sinc, L.orc_dphi_sinc):
            fn.restype = ctypes.c_double
            fn.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_bspline.restype = ctypes.c_double
        L.orc_bspline.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_rel_linf.restype = ctypes.c_double
        L.orc_rel_l2.restype = ctypes.c_double
        _lib = L
    return _lib


def _p(a, t=_f64p):
    return a.ctypes.data_as(t)


def _dims(N):
    N = np.ascontiguousarray(np.atleast_1d(np.asarray(N, dtype=np.int64)))
    return N, int(N.size)


def default_n(N):
    """``cunfft_init`` rule n_t = 2 * next_pow2(N_t) (src/cunfft.cu:296-299)."""
    out = []
    for v in np.atleast_1d(N):
        p = 1
        while p < int(v):
            p <<= 1
        out.append(2 * p)
    return np.array(out, dtype=np.int64)


def c2f(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)


class Drand48:
    """glibc drand48 stream (state 0 = an unseeded process)."""

    def __init__(self, state: int = 0):
        self.state = ctypes.c_uint64(state)

    def next(self) -> float:
        return lib().orc_drand48(ctypes.byref(self.state))

    def example_data(self, N, M):
        """getExampleData_uniDistr (src/cuda_util.cpp:151-160)."""
        N, d = _dims(N)
        Nt = int(np.prod(N))
        fhat = np.zeros(Nt, np.complex128)
        x = np.zeros((M, d), np.float64)
        lib().orc_example_data_unidistr(ctypes.byref(self.state), ctypes.c_int64(Nt),
                                        ctypes.c_int64(M), ctypes.c_int(d),
                                        _p(fhat.view(np.float64)), _p(x))
        return fhat, x

    def example_data_ad(self, M, d):
        """getExampleDataAd_uniDistr (src/cuda_util.cpp:162-171)."""
        f = np.zeros(M, np.complex128)
        x = np.zeros((M, d), np.float64)
        lib().orc_example_data_ad_unidistr(ctypes.byref(self.state), ctypes.c_int64(M),
                                           ctypes.c_int(d), _p(f.view(np.float64)), _p(x))
        return f, x


WINDOWS = {"gaussian": 0, "kaiser_bessel": 1, "sinc": 2}


def set_window(kind: str) -> None:
    """Window of the oracle's transforms: "gaussian" (the reference's),
    "kaiser_bessel" or "sinc" (NFFT3's).  Process-wide; tests restore "gaussian"."""
    lib().orc_set_window(WINDOWS[kind])


class window:
    """``with oracle.window("kaiser_bessel"): ...``"""

    def __init__(self, kind):
        self.kind = kind

    def __enter__(self):
        self.prev = lib().orc_get_window()
        set_window(self.kind)

    def __exit__(self, *exc):
        lib().orc_set_window(self.prev)


def kb_shape(N, n):
    return lib().orc_kb_shape(int(N), int(n))


def phi_kb(s, m, b):
    return lib().orc_phi_kb(float(s), int(m), float(b))


def dphi_kb(s, m, b):
    return lib().orc_dphi_kb(float(s), int(m), float(b))


def bessel_i0(z):
    return lib().orc_bessel_i0(float(z))


def sinc_shape(N, n, m):
    """a = (2 sigma - 1) / (2 sigma m) of the sinc-power window."""
    return lib().orc_sinc_shape(int(N), int(n), int(m))


def phi_sinc(s, m, a):
    return lib().orc_phi_sinc(float(s), int(m), float(a))


def dphi_sinc(s, m, a):
    return lib().orc_dphi_sinc(float(s), int(m), float(a))


def bspline(p, x):
    """Centred cardinal B-spline M_p(x) of order p."""
    return lib().orc_bspline(int(p), float(x))


def shape_b(N, n, m):
    return lib().orc_shape_b(int(N), int(n), int(m))


def window_start(n, x, m):
    return lib().orc_window_start(int(n), float(x), int(m))


def phi(t, b):
    return lib().orc_phi(float(t), float(b))


def rolloff_pad(N, n, m, fhat):
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    g = np.zeros(int(np.prod(n)), np.complex128)
    lib().orc_rolloff_pad(d, _p(N, _i64p), _p(n, _i64p), int(m), _p(c2f(fhat)),
                          _p(g.view(np.float64)))
    return g


def crop_rolloff(N, n, m, g):
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    fhat = np.zeros(int(np.prod(N)), np.complex128)
    lib().orc_crop_rolloff(d, _p(N, _i64p), _p(n, _i64p), int(m), _p(c2f(g)),
                           _p(fhat.view(np.float64)))
    return fhat


def fft(n, g, sign):
    n = np.ascontiguousarray(np.atleast_1d(n), dtype=np.int64)
    out = np.array(g, dtype=np.complex128, copy=True).reshape(-1)
    lib().orc_fft(int(n.size), _p(n, _i64p), _p(out.view(np.float64)), int(sign))
    return out


def gather(N, n, m, g, x):
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    f = np.zeros(M, np.complex128)
    lib().orc_gather(d, _p(N, _i64p), _p(n, _i64p), int(m), _p(c2f(g)), ctypes.c_int64(M),
                     _p(x), _p(f.view(np.float64)))
    return f


def scatter(N, n, m, x, f):
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    g = np.zeros(int(np.prod(n)), np.complex128)
    lib().orc_scatter(d, _p(N, _i64p), _p(n, _i64p), int(m), ctypes.c_int64(M), _p(x),
                      _p(c2f(f)), _p(g.view(np.float64)))
    return g


def nfft_trafo(N, n, m, x, fhat):
    """Forward NFFT f = B F D fhat, as cunfft_transform (src/cunfft.cu:196-207)."""
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    f = np.zeros(M, np.complex128)
    g = np.zeros(int(np.prod(n)), np.complex128)
    lib().orc_nfft_trafo(d, _p(N, _i64p), _p(n, _i64p), int(m), ctypes.c_int64(M), _p(x),
                         _p(c2f(fhat)), _p(f.view(np.float64)), _p(g.view(np.float64)))
    return f


def nfft_adjoint(N, n, m, x, f):
    """Adjoint NFFT, as cunfft_adjoint (src/cunfft.cu:209-224)."""
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    fhat = np.zeros(int(np.prod(N)), np.complex128)
    g = np.zeros(int(np.prod(n)), np.complex128)
    lib().orc_nfft_adjoint(d, _p(N, _i64p), _p(n, _i64p), int(m), ctypes.c_int64(M), _p(x),
                           _p(c2f(f)), _p(fhat.view(np.float64)), _p(g.view(np.float64)))
    return fhat


def ndft_trafo(N, x, fhat, rows=None, nthreads=0):
    """Direct NDFT f_j = sum_k fhat_k exp(-2 pi i k.x_j) (optionally only rows)."""
    N, d = _dims(N)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    if rows is not None:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nout = rows.size
    else:
        nout = M
    f = np.zeros(nout, np.complex128)
    lib().orc_ndft_trafo(d, _p(N, _i64p), ctypes.c_int64(M), _p(x), _p(c2f(fhat)),
                         _p(f.view(np.float64)), None if rows is None else _p(rows, _i64p),
                         ctypes.c_int64(nout), int(nthreads))
    return f


def ndft_adjoint(N, x, f, cols=None, nthreads=0):
    """Direct adjoint NDFT fhat_k = sum_j f_j exp(+2 pi i k.x_j) (optionally cols)."""
    N, d = _dims(N)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    Nt = int(np.prod(N))
    if cols is not None:
        cols = np.ascontiguousarray(cols, dtype=np.int64)
        nout = cols.size
    else:
        nout = Nt
    fhat = np.zeros(nout, np.complex128)
    lib().orc_ndft_adjoint(d, _p(N, _i64p), ctypes.c_int64(M), _p(x), _p(c2f(f)),
                           _p(fhat.view(np.float64)), None if cols is None else _p(cols, _i64p),
                           ctypes.c_int64(nout), int(nthreads))
    return fhat


def gather_dx(N, n, m, g, x):
    """d f_j / d x_jt of the implemented (truncated-window) gather, shape (M, d)."""
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    out = np.zeros((M, d), np.complex128)
    lib().orc_gather_dx(d, _p(N, _i64p), _p(n, _i64p), int(m), _p(c2f(g)), ctypes.c_int64(M),
                        _p(x), _p(out.view(np.float64)))
    return out


def nfft_dx(N, n, m, x, fhat):
    """d f_j / d x_jt of the forward NFFT: gather_dx on the forward grid."""
    N, d = _dims(N)
    n = np.ascontiguousarray(n, dtype=np.int64)
    g = fft(n, rolloff_pad(N, n, m, fhat), -1)
    return gather_dx(N, n, m, g, x)


def ndft_dx(N, x, fhat, rows=None, nthreads=0):
    """Analytic NDFT derivative, shape (len(rows) or M, d)."""
    N, d = _dims(N)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, d)
    M = x.shape[0]
    if rows is not None:
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        nout = rows.size
    else:
        nout = M
    out = np.zeros((nout, d), np.complex128)
    lib().orc_ndft_dx(d, _p(N, _i64p), ctypes.c_int64(M), _p(x), _p(c2f(fhat)),
                      _p(out.view(np.float64)), None if rows is None else _p(rows, _i64p),
                      ctypes.c_int64(nout), int(nthreads))
    return out


def rel_linf(a, b):
    """||a-b||_inf / ||a||_inf, the reference's error norm (src/cuda_util.cpp:237-258)."""
    a = np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.complex128).reshape(-1)
    return lib().orc_rel_linf(_p(a.view(np.float64)), _p(b.view(np.float64)),
                              ctypes.c_int64(a.size))


def rel_l2(a, b):
    a = np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.complex128).reshape(-1)
    return lib().orc_rel_l2(_p(a.view(np.float64)), _p(b.view(np.float64)),
                            ctypes.c_int64(a.size))
