"""Oversampled grids of n = 2^a 3^b cells (mixed-radix Stockham passes with
radix-3 stages, modular index wrap in the gridding): FP64 plans against the
oracle's NFFT to 1e-12 and FP32 to 1e-5 on FP32-rounded inputs, every
gridding configuration that fits, the FFT stage alone (unit impulses), and
a batch with sigma = 2 exactly for N_b in {96, 192}.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


CASES = [
    ([96], [192], 6, 3000),             # 1D, sigma = 2, radix 4,4,4,3
    ([1000], [1536], 4, 5000),          # 1D, 512 * 3
    ([96, 96], [192, 192], 4, 20000),   # 2D, sigma = 2
    ([64, 96], [128, 288], 4, 20000),   # 288 = 32 * 9: two radix-3 stages
    ([24, 32, 48], [48, 64, 96], 4, 20000),
    ([12, 16, 12], [24, 32, 24], 3, 5000),
]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_nonpow2_grid_matches_the_oracle(gpu, orc, case, precision):
    N, n, m, M = CASES[case]
    r = np.random.default_rng(60 + case)
    d = len(N)
    x = r.uniform(-0.5, 0.5, (M, d))
    if case % 2:   # clustered half
        c = r.normal(0.0, 0.1, (M // 2, d))
        x[M // 2:] = c - np.floor(c + 0.5)
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, M)
    ref_f = orc.nfft_trafo(N, n, m, x, q(fh, precision))
    ref_h = orc.nfft_adjoint(N, n, m, x, q(fv, precision))
    tol = 1e-12 if precision == "double" else 1e-5
    configs = [0] if d == 1 else list(range(gpu.TUNE_CONFIGS))
    with gpu.Plan(N, M, n=n, m=m, precision=precision) as p:
        for c in configs:
            try:
                p.config = c
            except gpu.CunfftError as e:
                # a tile that does not divide n or fit the LDS is refused
                print(N, n, precision, "config", c, "refused:", e)
                assert e.status == 2
                continue
            p.x[:] = x
            p.f_hat[:] = fh
            p.copyDataToDevice()
            p.transform()
            p.copyDataToHost()
            ef = orc.rel_l2(ref_f, p.f.astype(np.complex128))
            p.f[:] = fv
            p.copyDataToDeviceAd()
            p.adjoint()
            p.copyDataToHostAd()
            eh = orc.rel_l2(ref_h, p.f_hat.astype(np.complex128))
            print(N, n, precision, "config", c, ef, eh)
            assert ef <= tol and eh <= tol


@pytest.mark.parametrize("n", [[6], [24], [48], [96], [192], [384], [768], [1536], [3072],
                               [24, 96], [12, 18, 48]])
def test_fft_stage_of_unit_impulses(gpu, orc, n):
    """A single node at x = 0 with m = 1 touches few cells; comparing every
    entry of the GPU forward NFFT with the oracle for random fhat checks the
    FFT stage on each length (all frequencies pass through it)."""
    N = [v // 2 if (v // 2) % 2 == 0 else v // 2 + 1 for v in n]
    N = [min(v, nv) for v, nv in zip(N, n)]
    m = 1
    r = np.random.default_rng(7)
    x = r.uniform(-0.5, 0.5, (64, len(n)))
    fh = crandn(r, int(np.prod(N)))
    with gpu.Plan(N, 64, n=n, m=m) as p:
        p.x[:] = x
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, fh), p.f) <= 1e-12
        fv = crandn(r, 64)
        p.f[:] = fv
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()
        assert orc.rel_l2(orc.nfft_adjoint(N, n, m, x, fv), p.f_hat) <= 1e-12


def test_batch_with_sigma_two_exactly(gpu, orc):
    import torch

    from cunfft_amd.batch import Batch

    r = np.random.default_rng(96)
    sizes = [64, 96, 128, 192, 96, 192]
    N = np.array([[s, s] for s in sizes])
    n = 2 * N                                # sigma = 2 exactly: 192 and 384 for 96 and 192
    M = r.integers(1000, 4000, len(sizes))
    R, m = 2, 4
    xs = [r.uniform(-0.5, 0.5, (int(Mi), 2)) for Mi in M]
    fh = [np.stack([crandn(r, s * s) for _ in range(R)]) for s in sizes]
    fv = [np.stack([crandn(r, int(Mi)) for _ in range(R)]) for Mi in M]
    dev = torch.device("cuda")
    with Batch(N, M, m=m, n_rhs=R, n=n, precision="double") as b:
        for c in range(gpu.TUNE_CONFIGS):
            try:
                b.config = c
            except gpu.CunfftError as e:
                print("batch config", c, "refused:", e)
                assert e.status == 2
                continue
            for i in range(len(sizes)):
                xt = torch.from_numpy(xs[i]).to(dev)
                b.set_nodes_dev(i, xt.data_ptr())
                fi, _, ai, _ = b.tensors(i)
                fi.copy_(torch.from_numpy(fh[i]))
                ai.copy_(torch.from_numpy(fv[i]))
                torch.cuda.synchronize()
            b.execute()
            torch.cuda.synchronize()
            for i in range(len(sizes)):
                _, fo, _, ho = b.tensors(i)
                fo, ho = fo.cpu().numpy(), ho.cpu().numpy()
                for k in range(R):
                    Ni, ni = list(N[i]), list(n[i])
                    assert orc.rel_l2(orc.nfft_trafo(Ni, ni, m, xs[i], fh[i][k]), fo[k]) <= 1e-12
                    assert orc.rel_l2(orc.nfft_adjoint(Ni, ni, m, xs[i], fv[i][k]), ho[k]) <= 1e-12
            print("batch config", c, "ok")
