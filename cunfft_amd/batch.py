"""Batched NFFT over the C ABI (cunfft_batch_*): many uneven problems and
right-hand sides, every transform stage one launch over all of them, replayed
from a captured hipGraph.

The batch owns its device buffers (like a plan owns fhat_gpu / f_gpu,
src/cunfft.cu:262-280); :meth:`Batch.tensors` exposes them to torch without a
copy, through the CUDA array interface.  There is no arithmetic here and no
fallback: without the HIP library every call raises.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import TUNE_ADJOINT, TUNE_CONFIGS, TUNE_FORWARD, CunfftError, lib

_PROTOS = {
    "cunfft_batch_create": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                             ctypes.c_int], ctypes.c_int),
    "cunfft_batch_destroy": ([ctypes.c_void_p], None),
    "cunfft_batch_set_nodes_dev": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p], ctypes.c_int),
    "cunfft_batch_buffers": ([ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4, ctypes.c_int),
    "cunfft_batch_execute": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
    "cunfft_batch_autotune": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                               ctypes.c_void_p], ctypes.c_int),
    "cunfft_batch_get_config": ([ctypes.c_void_p, ctypes.c_int], ctypes.c_int),
    "cunfft_batch_set_config": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
}


_typed = []


def _lib():
    L = lib()
    if not _typed:
        for name, (args, res) in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _typed.append(True)
    return L


class _DevView:
    """A device buffer described by the CUDA array interface (no ownership)."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


class Batch:
    """P problems of one dimension d, cut-off m and precision; problem i has
    N[i] (d sizes), oversampled n[i] (None: 2 * next_pow2(N), as cunfft_init),
    M[i] nodes and n_rhs right-hand sides."""

    def __init__(self, N, M, m=4, n_rhs=1, n=None, precision="single"):
        N = np.asarray(N, dtype=np.int64)
        if N.ndim == 1:
            N = N[:, None]
        self.P, self.d = N.shape
        self.N = N
        self.M = np.asarray(M, dtype=np.int64).reshape(self.P)
        self.n_rhs = int(n_rhs)
        self.precision = precision
        self._prec = 64 if precision == "double" else 32
        if n is None:
            p2 = np.array([[1 << int(v - 1).bit_length() for v in row] for row in N], dtype=np.int64)
            self.n = 2 * p2
        else:
            self.n = np.asarray(n, dtype=np.int64).reshape(self.P, self.d)
        self._N_c = np.ascontiguousarray(N.ravel())
        self._n_c = np.ascontiguousarray(self.n.ravel())
        self._M_c = np.ascontiguousarray(self.M)
        h = ctypes.c_void_p()
        st = _lib().cunfft_batch_create(ctypes.byref(h), self._prec, self.d, self.P,
                                        self._N_c.ctypes.data, self._n_c.ctypes.data,
                                        self._M_c.ctypes.data, int(m), self.n_rhs)
        if st != 0:
            raise CunfftError(st, _lib().cunfft_last_message().decode(errors="replace"))
        self._h = h

    def _check(self, st):
        if st != 0:
            raise CunfftError(st, _lib().cunfft_last_message().decode(errors="replace"))

    def set_nodes_dev(self, problem, x_ptr, stream=None):
        self._check(_lib().cunfft_batch_set_nodes_dev(self._h, int(problem), ctypes.c_void_p(x_ptr),
                                                      ctypes.c_void_p(stream)))

    def buffers(self, problem):
        """Device pointers (fhat_in, f_out, f_in, fhat_out) of one problem."""
        p = [ctypes.c_void_p() for _ in range(4)]
        self._check(_lib().cunfft_batch_buffers(self._h, int(problem), *[ctypes.byref(v) for v in p]))
        return tuple(v.value for v in p)

    def tensors(self, problem):
        """torch views (n_rhs, N_total) / (n_rhs, M) of fhat_in, f_out, f_in, fhat_out."""
        import torch

        ts = "<c16" if self._prec == 64 else "<c8"
        Nt = int(np.prod(self.N[problem]))
        Mi = int(self.M[problem])
        out = []
        for ptr, n in zip(self.buffers(problem), (Nt, Mi, Mi, Nt)):
            t = torch.as_tensor(_DevView(ptr, self.n_rhs * n, ts), device="cuda")
            out.append(t.view(self.n_rhs, n))
        return tuple(out)

    def execute(self, directions=TUNE_FORWARD | TUNE_ADJOINT, stream=None):
        self._check(_lib().cunfft_batch_execute(self._h, int(directions), ctypes.c_void_p(stream)))

    def autotune(self, directions=TUNE_FORWARD | TUNE_ADJOINT, reps=5, stream=None):
        ms = (ctypes.c_float * (2 * TUNE_CONFIGS))()
        self._check(_lib().cunfft_batch_autotune(self._h, int(directions), int(reps),
                                                 ctypes.cast(ms, ctypes.c_void_p),
                                                 ctypes.c_void_p(stream)))
        v = [float(t) for t in ms]
        return self.config, v[:TUNE_CONFIGS], v[TUNE_CONFIGS:]

    @property
    def config(self):
        L = _lib()
        return (L.cunfft_batch_get_config(self._h, TUNE_FORWARD),
                L.cunfft_batch_get_config(self._h, TUNE_ADJOINT))

    @config.setter
    def config(self, c):
        pairs = [(TUNE_FORWARD | TUNE_ADJOINT, c)] if isinstance(c, int) else \
            [(TUNE_FORWARD, c[0]), (TUNE_ADJOINT, c[1])]
        for dirs, v in pairs:
            self._check(_lib().cunfft_batch_set_config(self._h, dirs, int(v)))

    def close(self):
        if getattr(self, "_h", None):
            _lib().cunfft_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
