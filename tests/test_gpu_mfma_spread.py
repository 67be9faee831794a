"""The 3D adjoint on the matrix cores (k_tile_spread_mfma, 8^3 tiles,
configuration 0), FP32 and FP64, against the oracle's NFFT adjoint (FP32: on
FP32-rounded inputs, to 1e-5 relative l2; FP64: to 1e-12):

* every window width it takes (m = 1..5, W = 4..12; m = 1..3 use one 16-plane
  set, m = 4, 5 two), uniform nodes;
* the Kaiser-Bessel and sinc-power windows (their weights are evaluated
  inline in the kernel);
* the second plane set alone: every node's window starting at tile-local
  plane 7 (rint(n x_0) - m = 7 mod 8), so each 4-node step runs on the
  shifted accumulators and the first set stays empty;
* a dense tile: 6000 nodes whose windows start in one tile (24 chunks of
  256, each 8 batches of 32) beside uniform ones;
* a node count that leaves partial batches and odd set sizes (the zero
  padding rows of each set);
* the same chunk sums as the VALU spread of configuration 1 (8x8x16 tiles):
  both within the tolerance of the oracle and of each other;
* a given binning gives a bit-identical grid call to call;
* 2D on 16^2 tiles (k_tile_spread_mfma2, configuration 1, FP32): m = 1..7, the
  Kaiser-Bessel and sinc windows, a dense tile with odd-sized last batches,
  agreement with the VALU spread of 32^2 tiles and bit-identical repeats; the
  batch's spread stage at m = 2, 6, 7 (matrix cores) and 8 (the VALU spread).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


@pytest.fixture(params=["single", "double"])
def prec(request):
    return request.param


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def rounded(a, prec):
    """The coefficients as the plan holds them (FP32-rounded for single)."""
    return a.astype(np.complex64).astype(np.complex128) if prec == "single" else a


def adjoint(gpu, prec, N, n, m, x, fv, flags=0, config=0, reps=1):
    with gpu.Plan(N, x.shape[0], n=n, m=m, precision=prec,
                  flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | flags) as p:
        p.config = config
        p.x[:] = x
        p.f[:] = fv
        p.copyDataToDeviceAd()
        out = []
        for _ in range(reps):
            p.adjoint()
            p.copyDataToHostAd()
            out.append(p.f_hat.astype(np.complex128).copy())
    return out if reps > 1 else out[0]


@pytest.mark.parametrize("m", [1, 2, 3, 4, 5])
def test_every_window_width_matches_the_oracle(gpu, orc, prec, m):
    N, n, M = [32, 32, 32], [64, 64, 64], 20000
    r = np.random.default_rng(500 + m)
    x = r.uniform(-0.5, 0.5, (M, 3))
    fv = crandn(r, M)
    h = adjoint(gpu, prec, N, n, m, x, fv)
    err = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, rounded(fv, prec)), h)
    print(f"{prec} m={m} W={2 * m + 2} rel-l2 vs oracle {err:.3e}")
    assert err <= TOL[prec]


@pytest.mark.parametrize("kind,flag", [("kaiser_bessel", "KAISER_BESSEL"), ("sinc", "SINC")])
def test_other_windows_match_the_oracle(gpu, orc, prec, kind, flag):
    N, n, m, M = [32, 32, 32], [64, 64, 64], 4, 20000
    r = np.random.default_rng(520)
    x = r.uniform(-0.5, 0.5, (M, 3))
    fv = crandn(r, M)
    h = adjoint(gpu, prec, N, n, m, x, fv, flags=getattr(gpu, flag))
    with orc.window(kind):
        ref = orc.nfft_adjoint(N, n, m, x, rounded(fv, prec))
    err = orc.rel_l2(ref, h)
    print(f"{prec} {kind} rel-l2 vs oracle {err:.3e}")
    assert err <= TOL[prec]


def starts_at_plane7(r, M, n, m):
    """x_0 in [0, 1/2) with rint(n x_0) - m = 8k + 7: window start at tile-local plane 7."""
    k = r.integers(0, int((n / 2 - 7 - m - 0.45) // 8) + 1, M)
    c = 8 * k + 7 + m + r.uniform(-0.45, 0.45, M)
    return c / n


def test_second_plane_set_alone(gpu, orc, prec):
    N, n, m, M = [32, 32, 32], [64, 64, 64], 4, 12000
    r = np.random.default_rng(530)
    x = r.uniform(-0.5, 0.5, (M, 3))
    x[:, 0] = starts_at_plane7(r, M, n[0], m)
    assert np.all([(orc.window_start(n[0], v, m) % n[0]) % 8 == 7 for v in x[:200, 0]])
    fv = crandn(r, M)
    h = adjoint(gpu, prec, N, n, m, x, fv)
    err = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, rounded(fv, prec)), h)
    print(f"{prec} all windows from plane 7: rel-l2 {err:.3e}")
    assert err <= TOL[prec]


def test_dense_tile_and_partial_batches(gpu, orc, prec):
    N, n, m = [32, 32, 32], [64, 64, 64], 4
    r = np.random.default_rng(540)
    # 6000 nodes whose windows start in one 8^3 grid tile (rint(n x) - m in [8, 16) on
    # every axis), all eight plane starts
    dense = (8 + m + r.uniform(-0.49, 7.49, (6000, 3))) / n[0]
    x = np.concatenate([dense, r.uniform(-0.5, 0.5, (4321, 3))])
    fv = crandn(r, x.shape[0])
    h = adjoint(gpu, prec, N, n, m, x, fv)
    err = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, rounded(fv, prec)), h)
    print(f"{prec} dense tile + 4321 uniform: rel-l2 {err:.3e}")
    assert err <= TOL[prec]


def test_agrees_with_the_valu_spread_and_repeats_bit_for_bit(gpu, orc, prec):
    N, n, m, M = [32, 32, 32], [64, 64, 64], 4, 30000
    r = np.random.default_rng(550)
    c = r.normal(0.0, 0.08, (M, 3))
    x = np.concatenate([c - np.floor(c + 0.5), r.uniform(-0.5, 0.5, (M, 3))])
    fv = crandn(r, x.shape[0])
    ha, hb = adjoint(gpu, prec, N, n, m, x, fv, config=0, reps=2)
    hv = adjoint(gpu, prec, N, n, m, x, fv, config=1)
    ref = orc.nfft_adjoint(N, n, m, x, rounded(fv, prec))
    ea, ev, eav = orc.rel_l2(ref, ha), orc.rel_l2(ref, hv), orc.rel_l2(hv, ha)
    print(f"{prec} matrix cores {ea:.3e}, VALU spread {ev:.3e}, between them {eav:.3e}")
    tol = TOL[prec]
    assert ea <= tol and ev <= tol and eav <= tol
    assert np.array_equal(ha, hb)


# ---- 2D on 16^2 tiles (configuration 1, FP32): one 32x32x2 product per two nodes

def test_2d_every_window_width_matches_the_oracle(gpu, orc):
    """m = 1..7 (W = 4..16: padded tiles of 19..31 rows, every one in one 32-row product)."""
    N, n, M = [64, 64], [128, 128], 30000
    for m in range(1, 8):
        r = np.random.default_rng(600 + m)
        x = r.uniform(-0.5, 0.5, (M, 2))
        fv = crandn(r, M)
        h = adjoint(gpu, "single", N, n, m, x, fv, config=1)
        err = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, rounded(fv, "single")), h)
        print(f"2D m={m} W={2 * m + 2} rel-l2 vs oracle {err:.3e}")
        assert err <= 1e-5


@pytest.mark.parametrize("kind,flag", [("kaiser_bessel", "KAISER_BESSEL"), ("sinc", "SINC")])
def test_2d_other_windows_match_the_oracle(gpu, orc, kind, flag):
    N, n, m, M = [64, 64], [128, 128], 4, 30000
    r = np.random.default_rng(620)
    x = r.uniform(-0.5, 0.5, (M, 2))
    fv = crandn(r, M)
    h = adjoint(gpu, "single", N, n, m, x, fv, flags=getattr(gpu, flag), config=1)
    with orc.window(kind):
        ref = orc.nfft_adjoint(N, n, m, x, rounded(fv, "single"))
    err = orc.rel_l2(ref, h)
    print(f"2D {kind} rel-l2 vs oracle {err:.3e}")
    assert err <= 1e-5


def test_2d_dense_tile_odd_batches_and_agreement_with_the_valu_spread(gpu, orc):
    """5001 nodes whose windows start in one 16^2 tile (several chunks, each batch of 32
    but the last odd-sized) beside uniform ones; against the oracle and the VALU spread
    of 32^2 tiles (configuration 0), and bit-identical call to call."""
    N, n, m = [64, 64], [128, 128], 4
    r = np.random.default_rng(640)
    dense = (16 + m + r.uniform(-0.49, 15.49, (5001, 2))) / n[0]
    x = np.concatenate([dense, r.uniform(-0.5, 0.5, (3333, 2))])
    fv = crandn(r, x.shape[0])
    ha, hb = adjoint(gpu, "single", N, n, m, x, fv, config=1, reps=2)
    hv = adjoint(gpu, "single", N, n, m, x, fv, config=0)
    ref = orc.nfft_adjoint(N, n, m, x, rounded(fv, "single"))
    ea, ev, eav = orc.rel_l2(ref, ha), orc.rel_l2(ref, hv), orc.rel_l2(hv, ha)
    print(f"2D matrix cores {ea:.3e}, VALU spread {ev:.3e}, between them {eav:.3e}")
    assert ea <= 1e-5 and ev <= 1e-5 and eav <= 1e-5
    assert np.array_equal(ha, hb)


@pytest.mark.parametrize("m", [2, 6, 7, 8])
def test_2d_batch_spread_other_window_widths(gpu, orc, m):
    """The batch's spread stage in configuration 1 (16^2 tiles): on the matrix cores
    for m <= 7 (W <= 16, padded tiles of at most 31 rows), the VALU spread at m = 8
    (33 rows); every (problem, right-hand side) against the oracle."""
    import torch

    from cunfft_amd.batch import Batch

    r = np.random.default_rng(660 + m)
    P, R = 5, 2
    N = np.array([[int(r.choice([16, 24, 32, 48, 64])) for _ in range(2)] for _ in range(P)])
    M = r.integers(500, 4000, P)
    xs = [r.uniform(-0.5, 0.5, (int(M[i]), 2)) for i in range(P)]
    fv = [crandn(r, R * int(M[i])).reshape(R, int(M[i])) for i in range(P)]
    dev = torch.device("cuda")
    with Batch(N, M, m=m, n_rhs=R, precision="single") as b:
        b.config = 1
        for i, x in enumerate(xs):
            xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
            b.set_nodes_dev(i, xt.data_ptr())
            _, _, f_in, _ = b.tensors(i)
            f_in.copy_(torch.from_numpy(fv[i]).to(f_in.dtype))
            torch.cuda.synchronize()
        b.execute(gpu.TUNE_ADJOINT)
        torch.cuda.synchronize()
        worst = 0.0
        for i in range(P):
            out = b.tensors(i)[3].cpu().numpy().astype(np.complex128)
            for k in range(R):
                e = orc.rel_l2(orc.nfft_adjoint(list(N[i]), list(b.n[i]), m, xs[i],
                                                rounded(fv[i][k], "single")), out[k])
                worst = max(worst, e)
                assert e <= 1e-5, (i, k, e)
    print(f"2D batch m={m} worst rel-l2 vs oracle {worst:.3e}")
