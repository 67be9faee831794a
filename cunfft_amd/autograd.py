"""Differentiable NFFT for torch: a thin torch.autograd.Function over the C ABI.

    f = A(x) fhat,   f_j = sum_k fhat_k exp(-2 pi i k.x_j)   (approximated by the NFFT)

is the forward NFFT of the reference (``cunfft_transform``, src/cunfft.cu:196-207).
Both gradients come from the library, on torch's current stream:

* with respect to fhat: the adjoint NFFT of the upstream gradient
  (``cunfft_adjoint``, src/cunfft.cu:209-224);
* with respect to the node coordinates x: the derivative-of-Gaussian gather
  gx[j, t] = Re(conj(g_j) * d f_j / d x_jt), whose window weights and their
  derivatives are recomputed in the backward pass (nothing is stored between
  forward and backward).

torch's convention for a real loss of complex tensors passes the conjugate
Wirtinger gradient scaled by two (for L = 0.5 ||f - y||^2 it is f - y); the
adjoint of a linear map is then the vector-Jacobian product, and the real
x-gradient is Re(conj(grad_f) df/dx).  Both entries are one C call,
``cunfft_backward_dev``.

This module is the test harness for the backward entry points; it holds no
arithmetic of its own and fails loudly without the HIP library.
"""
from __future__ import annotations

import numpy as np
import torch

from . import CUFFT_BOTH, CUFFT_INIT, NO_HOST_BUFFERS, Plan


class NFFTOperator:
    """One NFFT geometry (N, n, m, M, precision) on the GPU, usable with :func:`nfft`.

    The plan keeps device buffers only; the nodes are re-set from the tensor
    passed to each call (and re-binned on the plan's first transform after).
    """

    def __init__(self, N, M, n=None, m=4, precision="double", flags=0):
        """flags: extra plan flags (e.g. KAISER_BESSEL, COMPENSATED)."""
        self.plan = Plan(N, M, n=n, m=m, precision=precision,
                         flags=CUFFT_INIT | CUFFT_BOTH | NO_HOST_BUFFERS | flags)
        self.N = [int(v) for v in np.atleast_1d(N)]
        self.d = len(self.N)
        self.M = int(M)
        self.N_total = int(np.prod(self.N))
        self.cdtype = torch.complex128 if precision == "double" else torch.complex64
        self.rdtype = torch.float64 if precision == "double" else torch.float32

    @staticmethod
    def _stream():
        return torch.cuda.current_stream().cuda_stream

    def _check(self, fhat, x):
        if not (fhat.is_cuda and x.is_cuda):
            raise ValueError("nfft: fhat and x must be CUDA tensors")
        if fhat.dtype != self.cdtype or fhat.numel() != self.N_total:
            raise ValueError(f"nfft: fhat must be {self.cdtype} with {self.N_total} entries")
        if x.dtype != torch.float64 or tuple(x.shape) != (self.M, self.d):
            raise ValueError(f"nfft: x must be float64 of shape ({self.M}, {self.d})")

    def set_nodes(self, x):
        self.plan.set_nodes_dev(x.contiguous().data_ptr(), self._stream())

    def forward(self, fhat, x):
        self._check(fhat, x)
        self.set_nodes(x)
        f = torch.empty(self.M, dtype=self.cdtype, device=fhat.device)
        self.plan.transform_dev(fhat.contiguous().data_ptr(), f.data_ptr(), self._stream())
        return f

    def backward(self, fhat, x, grad_f, need_fhat=True, need_x=True):
        self.set_nodes(x)
        gf = grad_f.contiguous().to(self.cdtype)
        g_fhat = torch.empty(self.N_total, dtype=self.cdtype, device=gf.device) if need_fhat else None
        g_x = torch.empty((self.M, self.d), dtype=self.rdtype, device=gf.device) if need_x else None
        self.plan.backward_dev(fhat.contiguous().data_ptr() if need_x else 0, gf.data_ptr(),
                               g_fhat.data_ptr() if need_fhat else 0,
                               g_x.data_ptr() if need_x else 0, self._stream())
        return g_fhat, (g_x.to(torch.float64) if need_x else None)

    def close(self):
        self.plan.finalize()


class _NFFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fhat, x, op):
        ctx.op = op
        ctx.save_for_backward(fhat, x)
        return op.forward(fhat, x)

    @staticmethod
    def backward(ctx, grad_f):
        fhat, x = ctx.saved_tensors
        g_fhat, g_x = ctx.op.backward(fhat, x, grad_f, ctx.needs_input_grad[0],
                                      ctx.needs_input_grad[1])
        return g_fhat, g_x, None


def nfft(fhat: torch.Tensor, x: torch.Tensor, op: NFFTOperator) -> torch.Tensor:
    """f = A(x) fhat, differentiable in fhat (complex) and x (float64, shape (M, d))."""
    return _NFFT.apply(fhat, x, op)
