"""cunfft_adjoint_noalloc_dev: the adjoint NFFT with no device allocation.

Its node binning, sorted coefficients and partial-tile slots live in the
library's device-resident scratch array (scratch.hip), whose address is
taken once per device at plan set-up.  Checked through the C ABI:

* its output equals cunfft_adjoint_dev's on two calls in a row, the second
  with a smaller input (another plan, fewer nodes), and again on the first
  plan after the second (the scratch keeps the last call's contents): bit for
  bit for the plane-tiled kernels of d >= 2 (the same kernels on the same
  stable binning), to 1e-13 for d = 1, whose node-lane tiles add with atomics
  (the order of those additions, and so the last bits, may differ call to
  call); both match the oracle to 1e-12;
* calls after the first do not change the device's free memory (no
  allocation), while a first cunfft_adjoint_dev on a fresh plan does (its
  binning is allocated) -- so the measurement sees allocations;
* a plan whose binning does not fit the scratch is refused with
  CUNFFT_ERR_UNSUPPORTED and a message naming the sizes.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CASES = [([64, 64], 30000, 4), ([32, 32], 4000, 4), ([16, 16, 16], 12000, 3), ([1024], 5000, 6)]


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def setup(gpu, N, M, m, seed):
    r = np.random.default_rng(seed)
    dev = torch.device("cuda")
    x = torch.from_numpy(r.uniform(-0.5, 0.5, (M, len(N)))).to(dev)
    f = torch.from_numpy(crandn(r, M)).to(dev)
    p = gpu.Plan(N, M, m=m)
    p.set_nodes_dev(x.data_ptr())
    return p, x, f


def test_noalloc_adjoint_equals_the_adjoint_entry_point(gpu, orc):
    dev = torch.device("cuda")
    plans = [setup(gpu, N, M, m, 50 + k) for k, (N, M, m) in enumerate(CASES)]
    try:
        outs = []
        for (p, x, f), (N, M, m) in zip(plans, CASES):
            Nt = int(np.prod(N))
            h_na = torch.full((Nt,), complex("nan"), dtype=torch.complex128, device=dev)
            p.adjoint_noalloc_dev(f.data_ptr(), h_na.data_ptr())
            h_ref = torch.zeros(Nt, dtype=torch.complex128, device=dev)
            p.adjoint_dev(f.data_ptr(), h_ref.data_ptr())
            torch.cuda.synchronize()
            if len(N) >= 2:
                assert torch.equal(h_na, h_ref), (N, M)
            else:
                assert float((h_na - h_ref).norm() / h_ref.norm()) <= 1e-13, (N, M)
            ref = orc.nfft_adjoint(N, orc.default_n(N).tolist(), m, x.cpu().numpy(), f.cpu().numpy())
            assert orc.rel_l2(ref, h_na.cpu().numpy()) <= 1e-12
            outs.append(h_ref)
        # the first plan again, after the smaller ones left their data in the scratch
        p, x, f = plans[0]
        h = torch.zeros_like(outs[0])
        p.adjoint_noalloc_dev(f.data_ptr(), h.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(h, outs[0])   # 2D: plane-tiled, bit-reproducible
    finally:
        for p, _, _ in plans:
            p.finalize()


def test_noalloc_adjoint_allocates_nothing(gpu):
    dev = torch.device("cuda")
    N, M, m = [64, 64], 30000, 4
    a = setup(gpu, N, M, m, 1)
    b = setup(gpu, [32, 32], 4000, m, 2)
    c = setup(gpu, N, M, m, 3)
    try:
        h = torch.zeros(64 * 64, dtype=torch.complex128, device=dev)
        hb = torch.zeros(32 * 32, dtype=torch.complex128, device=dev)
        a[0].adjoint_noalloc_dev(a[2].data_ptr(), h.data_ptr())   # loads the kernels
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        for _ in range(3):
            b[0].adjoint_noalloc_dev(b[2].data_ptr(), hb.data_ptr())
            a[0].adjoint_noalloc_dev(a[2].data_ptr(), h.data_ptr())
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert free1 == free0, (free0, free1)
        # the allocating entry point's first call on a fresh plan is seen
        c[0].adjoint_dev(c[2].data_ptr(), h.data_ptr())
        torch.cuda.synchronize()
        free2, _ = torch.cuda.mem_get_info()
        assert free2 < free1
    finally:
        for p, _, _ in (a, b, c):
            p.finalize()


def test_noalloc_adjoint_refuses_what_does_not_fit(gpu):
    dev = torch.device("cuda")
    N, M = [128, 128, 128], 1 << 21
    p, x, f = setup(gpu, N, M, 4, 9)
    try:
        h = torch.zeros(128 ** 3, dtype=torch.complex128, device=dev)
        with pytest.raises(gpu.CunfftError) as e:
            p.adjoint_noalloc_dev(f.data_ptr(), h.data_ptr())
        assert e.value.status == 2 and "MiB" in str(e.value)
    finally:
        p.finalize()
