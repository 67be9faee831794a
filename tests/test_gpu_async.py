"""Asynchronous copy boundary (CUNFFT_ASYNC_COPY): copies and transforms of
two plans queued back to back on their own streams, with the results read
after the copies to the host; both against the oracle's NFFT."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_two_async_plans_interleaved(gpu, orc):
    r = np.random.default_rng(12)
    specs = [([48, 64], [128, 128], 4, 30000, "double"), ([16, 24, 32], [32, 64, 64], 3, 20000, "single")]
    plans, data = [], []
    for N, n, m, M, prec in specs:
        p = gpu.Plan(N, M, n=n, m=m, precision=prec,
                     flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.ASYNC_COPY)
        x = r.uniform(-0.5, 0.5, (M, len(N)))
        fh = r.standard_normal(int(np.prod(N))) + 1j * r.standard_normal(int(np.prod(N)))
        fv = r.standard_normal(M) + 1j * r.standard_normal(M)
        plans.append(p)
        data.append((N, n, m, x, fh, fv, prec))
    try:
        for p, (N, n, m, x, fh, fv, prec) in zip(plans, data):
            p.x[:] = x
            p.f_hat[:] = fh
            p.copyDataToDevice()          # returns with the copy queued
        for p in plans:
            p.transform()                 # queued behind the copy
        outs = []
        for p in plans:
            p.copyDataToHost()            # returns with the data
            outs.append(p.f.astype(np.complex128).copy())
        for p, (N, n, m, x, fh, fv, prec) in zip(plans, data):
            p.synchronize()
            p.f[:] = fv
            p.copyDataToDeviceAd()
            p.adjoint()
        adj = []
        for p in plans:
            p.copyDataToHostAd()
            adj.append(p.f_hat.astype(np.complex128).copy())
        for (N, n, m, x, fh, fv, prec), f, h in zip(data, outs, adj):
            q = (lambda a: a.astype(np.complex64).astype(np.complex128)) if prec == "single" else (lambda a: a)
            tol = 1e-12 if prec == "double" else 1e-5
            ef = orc.rel_l2(orc.nfft_trafo(N, n, m, x, q(fh)), f)
            eh = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, q(fv)), h)
            print(N, prec, "async", ef, eh)
            assert ef <= tol and eh <= tol
    finally:
        for p in plans:
            p.finalize()


def test_async_plan_sequence_on_one_stream(gpu, orc):
    """Several transforms of one plan, each copy-in queued behind the previous
    copy-out: stream order keeps every result."""
    N, n, m, M = [32, 32], [64, 64], 4, 5000
    r = np.random.default_rng(13)
    x = r.uniform(-0.5, 0.5, (M, 2))
    with gpu.Plan(N, M, n=n, m=m, flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.ASYNC_COPY) as p:
        for k in range(4):
            fh = r.standard_normal(1024) + 1j * r.standard_normal(1024)
            p.x[:] = x
            p.f_hat[:] = fh
            p.copyDataToDevice()
            p.transform()
            p.copyDataToHost()
            assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, fh), p.f) <= 1e-12
