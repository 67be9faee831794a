"""Batched transforms (cunfft_batch_*): every (problem, right-hand side) of a
batch against the oracle's NFFT, for each gridding configuration, FP64 to
1e-12 and FP32 to 1e-5 (on FP32-rounded inputs); the batch against single
plans of the same problems; re-binning when one problem's nodes change; the
batch autotuner.  Problems are uneven on purpose: different N, n (default
rule n = 2 next_pow2(N) and explicit), M and node distributions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(r, *shape):
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


def make_problems(r, d, P):
    sizes = {1: [128, 200, 256, 300], 2: [16, 24, 32, 48, 64], 3: [8, 12, 16]}[d]
    N = np.array([[int(r.choice(sizes)) for _ in range(d)] for _ in range(P)])
    M = r.integers(300, 3000, P)
    xs = []
    for i in range(P):
        if i % 3 == 2:   # clustered
            c = r.normal(0.0, 0.05, (M[i], d))
            xs.append(c - np.floor(c + 0.5))
        else:
            xs.append(r.uniform(-0.5, 0.5, (M[i], d)))
    return N, M, xs


def fill_and_run(gpu_batch, b, xs, fh, fv, torch, dev):
    for i, x in enumerate(xs):
        xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        b.set_nodes_dev(i, xt.data_ptr())
        fhat_in, _, f_in, _ = b.tensors(i)
        fhat_in.copy_(torch.from_numpy(fh[i]).to(fhat_in.dtype))
        f_in.copy_(torch.from_numpy(fv[i]).to(f_in.dtype))
        torch.cuda.synchronize()   # xt is released after the node copy has run
    b.execute()
    torch.cuda.synchronize()
    out_f = [b.tensors(i)[1].cpu().numpy().astype(np.complex128) for i in range(len(xs))]
    out_h = [b.tensors(i)[3].cpu().numpy().astype(np.complex128) for i in range(len(xs))]
    return out_f, out_h


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_batch_matches_the_oracle_for_every_configuration(gpu, orc, precision, d):
    import torch

    from cunfft_amd.batch import Batch

    r = np.random.default_rng(500 + d)
    P, R, m = 7, 3, 4 if d > 1 else 5
    N, M, xs = make_problems(r, d, P)
    fh = [crandn(r, R, int(np.prod(N[i]))) for i in range(P)]
    fv = [crandn(r, R, int(M[i])) for i in range(P)]
    tol = 1e-12 if precision == "double" else 1e-5
    dev = torch.device("cuda")
    configs = [0] if d == 1 else list(range(gpu.TUNE_CONFIGS))
    with Batch(N, M, m=m, n_rhs=R, precision=precision) as b:
        for c in configs:
            try:
                b.config = c
            except gpu.CunfftError as e:
                print(d, precision, "config", c, "refused:", e)
                assert e.status == 2 and 0 < c < gpu.TILE_SHAPES
                continue
            assert b.config == (c, c)
            out_f, out_h = fill_and_run(gpu, b, xs, fh, fv, torch, dev)
            worst = 0.0
            for i in range(P):
                Ni, ni = list(N[i]), list(b.n[i])
                for k in range(R):
                    ef = orc.rel_l2(orc.nfft_trafo(Ni, ni, m, xs[i], q(fh[i][k], precision)),
                                    out_f[i][k])
                    eh = orc.rel_l2(orc.nfft_adjoint(Ni, ni, m, xs[i], q(fv[i][k], precision)),
                                    out_h[i][k])
                    worst = max(worst, ef, eh)
                    assert ef <= tol and eh <= tol, (c, i, k, ef, eh)
            print(d, precision, "config", c, "worst rel-l2 vs oracle", worst)


def test_batch_equals_single_plans_and_rebins_changed_nodes(gpu, orc):
    """Same kernels and binning as single plans (forward bit-identical); after
    new nodes for one problem, that problem is re-binned and the rest kept."""
    import torch

    from cunfft_amd.batch import Batch

    r = np.random.default_rng(77)
    P, R, m = 5, 2, 4
    N, M, xs = make_problems(r, 2, P)
    fh = [crandn(r, R, int(np.prod(N[i]))) for i in range(P)]
    fv = [crandn(r, R, int(M[i])) for i in range(P)]
    dev = torch.device("cuda")
    with Batch(N, M, m=m, n_rhs=R, precision="double") as b:
        out_f, out_h = fill_and_run(gpu, b, xs, fh, fv, torch, dev)
        for i in range(P):
            with gpu.Plan(list(N[i]), int(M[i]), n=list(b.n[i]), m=m) as p:
                p.x[:] = xs[i]
                for k in range(R):
                    p.f_hat[:] = fh[i][k]
                    p.copyDataToDevice()
                    p.transform()
                    p.copyDataToHost()
                    assert np.array_equal(p.f, out_f[i][k])
                    p.f[:] = fv[i][k]
                    p.copyDataToDeviceAd()
                    p.adjoint()
                    p.copyDataToHostAd()
                    np.testing.assert_allclose(p.f_hat, out_h[i][k], rtol=0,
                                               atol=1e-12 * np.abs(out_h[i][k]).max())
        # new nodes for problem 3 only
        xs2 = list(xs)
        xs2[3] = r.uniform(-0.5, 0.5, xs[3].shape)
        x3 = torch.from_numpy(xs2[3]).to(dev)
        b.set_nodes_dev(3, x3.data_ptr())
        b.execute()
        torch.cuda.synchronize()
        for i in range(P):
            f_i = b.tensors(i)[1].cpu().numpy()
            h_i = b.tensors(i)[3].cpu().numpy()
            for k in range(R):
                assert orc.rel_l2(orc.nfft_trafo(list(N[i]), list(b.n[i]), m, xs2[i], fh[i][k]),
                                  f_i[k]) <= 1e-12
                assert orc.rel_l2(orc.nfft_adjoint(list(N[i]), list(b.n[i]), m, xs2[i], fv[i][k]),
                                  h_i[k]) <= 1e-12


def test_batch_forward_and_adjoint_alone_and_autotune(gpu, orc):
    import torch

    from cunfft_amd.batch import Batch

    r = np.random.default_rng(91)
    P, R, m = 6, 2, 4
    N, M, xs = make_problems(r, 2, P)
    fh = [crandn(r, R, int(np.prod(N[i]))) for i in range(P)]
    fv = [crandn(r, R, int(M[i])) for i in range(P)]
    dev = torch.device("cuda")
    with Batch(N, M, m=m, n_rhs=R, precision="single") as b:
        fill_and_run(gpu, b, xs, fh, fv, torch, dev)
        chosen, ms_f, ms_a = b.autotune(reps=3)
        print("batch autotune", chosen, ms_f, ms_a)
        for dd, ms in enumerate((ms_f, ms_a)):
            valid = [c for c in range(gpu.TUNE_CONFIGS) if ms[c] >= 0]
            assert 0 in valid and chosen[dd] in valid
            assert ms[chosen[dd]] == min(ms[c] for c in valid)
        assert b.config == chosen
        # forward alone, then adjoint alone, on fresh outputs
        for i in range(P):
            b.tensors(i)[1].zero_()
            b.tensors(i)[3].zero_()
        b.execute(gpu.TUNE_FORWARD)
        torch.cuda.synchronize()
        assert all(float(b.tensors(i)[3].abs().max()) == 0.0 for i in range(P))
        b.execute(gpu.TUNE_ADJOINT)
        torch.cuda.synchronize()
        for i in range(P):
            f_i, h_i = b.tensors(i)[1].cpu().numpy(), b.tensors(i)[3].cpu().numpy()
            for k in range(R):
                Ni, ni = list(N[i]), list(b.n[i])
                assert orc.rel_l2(orc.nfft_trafo(Ni, ni, m, xs[i], q(fh[i][k], "single")),
                                  f_i[k].astype(np.complex128)) <= 1e-5
                assert orc.rel_l2(orc.nfft_adjoint(Ni, ni, m, xs[i], q(fv[i][k], "single")),
                                  h_i[k].astype(np.complex128)) <= 1e-5


def test_batch_rejects_bad_arguments(gpu):
    from cunfft_amd.batch import Batch

    with pytest.raises(gpu.CunfftError):
        Batch([[30, 32]], [100], m=4, n=[[40, 64]])   # n not of the form 2^a 3^b
    with pytest.raises(gpu.CunfftError):
        Batch([[31, 32]], [100], m=4)                  # odd N
    with pytest.raises(gpu.CunfftError):
        Batch([[64]], [100], m=4)                      # n = 128 < the 1D tile (256)
    with Batch([[32, 32]], [100], m=4) as b:
        with pytest.raises(gpu.CunfftError):
            b.execute(4)
        with pytest.raises(gpu.CunfftError):
            b.config = 7
