"""Axes longer than one FFT pass (n_t > 4096): the split (four-step) FFT.

The reference plans cuFFT on any n (src/cunfft_wrapper.cu:265-286) and its
recorded 1D runs go to N = 2^22 under cunfft_init's n = 2 next_pow2(N)
(Tests/Runtime/test_m4.txt:10-19).  Here an axis n_t = L1 L2 > 4096 runs as
two LDS-resident passes (L2-point FFTs in place, then L1-point FFTs with the
w_n twiddles written in natural order to a second grid); the roll-off, pad
and crop use the full axis index.  Every case is checked against the oracle:
FP64 to 1e-12, FP32 to 1e-5 on FP32-rounded inputs.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


# (N, n or None for cunfft_init's rule, m, M)
CASES = [
    ([8192], None, 4, 20000),          # n = 2^14 = 128 x 128: the first length past one pass
    ([16384], None, 6, 20000),         # n = 2^15
    ([3072], [6144], 4, 9000),         # n = 2^11 3, non-power-of-two split (64 x 96)
    ([262144], None, 4, 50000),        # n = 2^19
    ([4194304], None, 4, 100000),      # N = 2^22, n = 2^23: the largest recorded reference run
    ([4096, 16], None, 4, 20000),      # split first axis, strided (S = 32)
    ([16, 3072], [32, 6144], 4, 20000),  # split last axis: from / to fhat, non-power-of-two
    ([8, 4096, 8], None, 3, 20000),    # split middle axis
]


def run(gpu, N, n, m, x, fh, fv, precision, flags=0):
    with gpu.Plan(N, x.shape[0], n=n, m=m, precision=precision,
                  flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | flags) as p:
        p.x[:] = x
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        f = p.f.astype(np.complex128).copy()
        p.copy_g_ToHost()
        g = p.g.astype(np.complex128).copy()
        p.f[:] = fv
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()
        return f, g, p.f_hat.astype(np.complex128).copy()


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_split_axis_matches_the_oracle(gpu, orc, case, precision):
    N, n, m, M = CASES[case]
    d = len(N)
    n = n or orc.default_n(N).tolist()
    r = np.random.default_rng(300 + case)
    x = r.uniform(-0.5, 0.5, (M, d))
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, M)
    print(f"case {N} n={n} m={m} M={M} {precision}", flush=True)
    f, g, h = run(gpu, N, n, m, x, fh, fv, precision)
    ref_g = orc.fft(n, orc.rolloff_pad(N, n, m, q(fh, precision)), -1)
    ref_f = orc.nfft_trafo(N, n, m, x, q(fh, precision))
    ref_h = orc.nfft_adjoint(N, n, m, x, q(fv, precision))
    eg, ef, eh = orc.rel_l2(ref_g, g), orc.rel_l2(ref_f, f), orc.rel_l2(ref_h, h)
    print(f"  grid {eg:.2e} forward {ef:.2e} adjoint {eh:.2e}")
    tol = 1e-12 if precision == "double" else 1e-5
    assert eg <= tol and ef <= tol and eh <= tol


def test_split_axis_untiled_kernels_and_gradient(gpu, orc):
    N, m, M = [8192], 4, 5000
    n = orc.default_n(N).tolist()
    r = np.random.default_rng(7)
    x = r.uniform(-0.5, 0.5, (M, 1))
    fh, fv = crandn(r, N[0]), crandn(r, M)
    f, _, h = run(gpu, N, None, m, x, fh, fv, "double", gpu.PLAIN_GATHER | gpu.ATOMIC_SCATTER)
    assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, fh), f) <= 1e-12
    assert orc.rel_l2(orc.nfft_adjoint(N, n, m, x, fv), h) <= 1e-12
    dev = torch.device("cuda")
    with gpu.Plan(N, M, m=m) as p:
        # the device arrays stay referenced until the kernels have run
        xt, fht, fvt = torch.from_numpy(x).to(dev), torch.from_numpy(fh).to(dev), torch.from_numpy(fv).to(dev)
        gx = torch.zeros((M, 1), dtype=torch.float64, device=dev)
        p.set_nodes_dev(xt.data_ptr())
        p.grad_x_dev(fht.data_ptr(), fvt.data_ptr(), gx.data_ptr())
        torch.cuda.synchronize()
    ref = (np.conj(fv)[:, None] * orc.nfft_dx(N, n, m, x, fh)).real
    assert np.linalg.norm(gx.cpu().numpy() - ref) / np.linalg.norm(ref) <= 1e-12


def test_cunfft_init_takes_the_recorded_1d_sizes(gpu):
    """cunfft_init (default m, n = 2 next_pow2(N)) for the reference's recorded
    1D sizes l_N = 13..22 (Tests/Runtime/test_m4.txt:10-19) is accepted."""
    for l in range(13, 23):
        with gpu.Plan([1 << l], 1000) as p:
            assert p.n == [1 << (l + 1)]
    with pytest.raises(gpu.CunfftError) as e:
        gpu.Plan([1 << 24], 10)   # n = 2^25 exceeds CUNFFT_MAX_AXIS
    assert e.value.status == 2
