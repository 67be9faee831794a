"""Sinc-power window (CUNFFT_SINC): GPU plans against the oracle's sinc NFFT
(FP64 to 1e-12, FP32 to 1e-5 on FP32-rounded inputs) in every gridding
configuration and on the plain kernels, against the NDFT (the window's
accuracy: ~1e-5 at m = 4, ~1e-10 at m = 8; the Gaussian gives ~1e-4 and
~1e-8), the x-gradient against the oracle's derivative of the same window, and
the argument checks (n_t > N_t, not together with Kaiser-Bessel).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


def run(gpu, N, n, m, x, fh, fv, precision, flags=0, config=None):
    with gpu.Plan(N, x.shape[0], n=n, m=m, precision=precision,
                  flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.SINC | flags) as p:
        if config is not None:
            p.config = config
        p.x[:] = x
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        f = p.f.astype(np.complex128).copy()
        p.f[:] = fv
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()
        return f, p.f_hat.astype(np.complex128).copy()


CASES = [([128], [256], 6, 3000), ([48, 64], [96, 128], 4, 20000), ([24, 16, 32], [48, 32, 64], 4, 20000),
         ([96, 64], [192, 128], 5, 8000)]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_sinc_matches_the_oracle(gpu, orc, case, precision):
    N, n, m, M = CASES[case]
    d = len(N)
    r = np.random.default_rng(180 + case)
    x = r.uniform(-0.5, 0.5, (M, d))
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, M)
    with orc.window("sinc"):
        ref_f = orc.nfft_trafo(N, n, m, x, q(fh, precision))
        ref_h = orc.nfft_adjoint(N, n, m, x, q(fv, precision))
    tol = 1e-12 if precision == "double" else 1e-5
    variants = [dict(config=c) for c in ([0] if d == 1 else range(gpu.TUNE_CONFIGS))]
    variants.append(dict(flags=gpu.PLAIN_GATHER | gpu.ATOMIC_SCATTER))
    for v in variants:
        try:
            f, h = run(gpu, N, n, m, x, fh, fv, precision, **v)
        except gpu.CunfftError as e:
            print(N, precision, v, "refused:", e)
            assert e.status == 2
            continue
        ef, eh = orc.rel_l2(ref_f, f), orc.rel_l2(ref_h, h)
        print(N, precision, v, "vs oracle", ef, eh)
        assert ef <= tol and eh <= tol


@pytest.mark.parametrize("m,band", [(4, 2.5e-5), (6, 1e-7), (8, 4e-10)])
def test_sinc_against_the_ndft(gpu, orc, m, band):
    N, n, M = [32, 48], [64, 96], 4000
    r = np.random.default_rng(190 + m)
    x = r.uniform(-0.5, 0.5, (M, 2))
    fh = crandn(r, 32 * 48)
    fv = crandn(r, M)
    f, h = run(gpu, N, n, m, x, fh, fv, "double")
    ef = orc.rel_l2(orc.ndft_trafo(N, x, fh), f)
    eh = orc.rel_l2(orc.ndft_adjoint(N, x, fv), h)
    print("sinc m", m, "vs NDFT", ef, eh)
    assert ef <= band and eh <= band


def test_sinc_gradient(gpu, orc):
    import torch

    from cunfft_amd.autograd import NFFTOperator

    N, n, m, M = [32, 32], [64, 64], 4, 3000
    r = np.random.default_rng(199)
    x = r.uniform(-0.5, 0.5, (M, 2))
    fh = crandn(r, 1024)
    v = crandn(r, M)
    op = NFFTOperator(N, M, n=n, m=m, precision="double", flags=gpu.SINC)
    try:
        dev = torch.device("cuda")
        g_fhat, g_x = op.backward(torch.from_numpy(fh).to(dev), torch.from_numpy(x).to(dev),
                                  torch.from_numpy(v).to(dev))
        torch.cuda.synchronize()
        with orc.window("sinc"):
            dfdx = orc.nfft_dx(N, n, m, x, fh)
            ref_h = orc.nfft_adjoint(N, n, m, x, v)
        ref_gx = (np.conj(v)[:, None] * dfdx).real
        ex = np.linalg.norm(g_x.cpu().numpy() - ref_gx) / np.linalg.norm(ref_gx)
        eh = orc.rel_l2(ref_h, g_fhat.cpu().numpy())
        print("sinc gradients: x", ex, "fhat", eh)
        assert ex <= 1e-10 and eh <= 1e-12
    finally:
        op.close()


def test_sinc_argument_checks(gpu):
    with pytest.raises(gpu.CunfftError) as e:
        gpu.Plan([32, 32], 100, n=[32, 64], m=4, flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.SINC)
    assert e.value.status == 1
    with pytest.raises(gpu.CunfftError) as e:
        gpu.Plan([32, 32], 100, m=4,
                 flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.SINC | gpu.KAISER_BESSEL)
    assert e.value.status == 1
