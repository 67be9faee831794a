"""Nodes outside [-1/2, 1/2) and non-finite nodes on the GPU.

The reference takes any node x: it masks the window indices with n - 1
(src/cunfft_kernel.cu:528-531), i.e. it uses u mod n with
u = llrint(n x) - signbit(x) - m.  The gridding kernels here reduce the
window start mod n the same way (window_cell), so nodes in [-3/2, 3/2)
match the oracle, which wraps with the same modulo, in every gridding
configuration and for the x-gradient.  NaN / inf / huge nodes must not
produce an index outside the grid: the calls succeed and the finite nodes'
forward values are unaffected.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CASES = [([128], 6), ([32, 48], 4), ([16, 16, 16], 3)]


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def configs(gpu, d):
    """(config, flags) pairs: every tiled configuration and the untiled kernels."""
    out = [(c, 0) for c in range(gpu.TUNE_CONFIGS if d >= 2 else 1)]
    out.append((None, gpu.PLAIN_GATHER | gpu.ATOMIC_SCATTER))
    return out


def run(gpu, N, m, x, fh, fv, config, flags, precision="double"):
    with gpu.Plan(N, x.shape[0], m=m, precision=precision,
                  flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | flags) as p:
        if config is not None:
            try:
                p.config = config
            except gpu.CunfftError as e:   # a tile shape that does not fit this grid
                assert e.status == 2, e
                return None, None
        p.x[:] = x
        p.f_hat[:] = fh
        p.f[:] = fv
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        f = p.f.astype(np.complex128).copy()
        p.f[:] = fv
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()
        return f, p.f_hat.astype(np.complex128).copy()


@pytest.mark.parametrize("N,m", CASES)
def test_nodes_outside_the_unit_box_match_the_oracle(gpu, orc, N, m):
    r = np.random.default_rng(11 + len(N))
    d, M = len(N), 3000
    n = orc.default_n(N).tolist()
    x = r.uniform(-1.5, 1.5, (M, d))
    x[:8] = np.array([-1.5, -1.0, -0.5, 0.5, 1.0, 1.25, -0.0, 0.0])[:, None]
    fh, fv = crandn(r, int(np.prod(N))), crandn(r, M)
    ref_f = orc.nfft_trafo(N, n, m, x, fh)
    ref_h = orc.nfft_adjoint(N, n, m, x, fv)
    for config, flags in configs(gpu, d):
        f, h = run(gpu, N, m, x, fh, fv, config, flags)
        if f is None:
            continue
        ef, eh = orc.rel_l2(ref_f, f), orc.rel_l2(ref_h, h)
        print(N, "config", config, "flags", flags, "fwd", ef, "adj", eh)
        assert ef <= 1e-12 and eh <= 1e-12


@pytest.mark.parametrize("N,m", CASES)
def test_gradient_for_nodes_outside_the_unit_box(gpu, orc, N, m):
    r = np.random.default_rng(5 + len(N))
    d, M = len(N), 2000
    n = orc.default_n(N).tolist()
    x = r.uniform(-1.5, 1.5, (M, d))
    fh, v = crandn(r, int(np.prod(N))), crandn(r, M)
    dfdx = orc.nfft_dx(N, n, m, x, fh).reshape(M, d)
    ref = (np.conj(v)[:, None] * dfdx).real
    dev = torch.device("cuda")
    with gpu.Plan(N, M, m=m) as p:
        xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        fht = torch.from_numpy(fh).to(dev)
        vt = torch.from_numpy(v).to(dev)
        gx = torch.zeros((M, d), dtype=torch.float64, device=dev)
        p.set_nodes_dev(xt.data_ptr())
        p.grad_x_dev(fht.data_ptr(), vt.data_ptr(), gx.data_ptr())
        torch.cuda.synchronize()
    err = np.linalg.norm(gx.cpu().numpy() - ref) / np.linalg.norm(ref)
    print(N, "x-gradient rel-l2 vs oracle", err)
    assert err <= 1e-12


@pytest.mark.parametrize("N,m", CASES)
def test_non_finite_nodes_stay_in_the_grid(gpu, orc, N, m):
    r = np.random.default_rng(23)
    d, M = len(N), 2000
    n = orc.default_n(N).tolist()
    x = r.uniform(-0.5, 0.5, (M, d))
    bad = np.array([3, 100, 777, 1500, 1999])
    x[bad, 0] = [np.nan, np.inf, -np.inf, 1e300, -2.0 ** 60]
    good = np.setdiff1d(np.arange(M), bad)
    fh, fv = crandn(r, int(np.prod(N))), crandn(r, M)
    ref_f = orc.nfft_trafo(N, n, m, x[good], fh)
    for config, flags in configs(gpu, d):
        f, h = run(gpu, N, m, x, fh, fv, config, flags)
        if f is None:
            continue
        ef = orc.rel_l2(ref_f, f[good])
        print(N, "config", config, "flags", flags, "finite nodes fwd", ef)
        assert ef <= 1e-12
        assert h.shape == (int(np.prod(N)),)


def test_ndft_output_index_out_of_range_gives_nan(gpu, orc):
    """cundft_*_dev with listed outputs: an index outside the outputs yields
    NaN there (no read past the arrays); the valid ones match the oracle."""
    r = np.random.default_rng(4)
    N, M = [16, 16], 300
    x = r.uniform(-0.5, 0.5, (M, 2))
    fh, fv = crandn(r, 256), crandn(r, M)
    dev = torch.device("cuda")
    rows = np.array([0, 5, M - 1, M, -1, 10 ** 12, 7], dtype=np.int64)
    cols = np.array([0, 255, 256, -3, 2 ** 40, 100], dtype=np.int64)
    with gpu.Plan(N, M) as p:
        xt = torch.from_numpy(x).to(dev)
        p.set_nodes_dev(xt.data_ptr())
        fht, fvt = torch.from_numpy(fh).to(dev), torch.from_numpy(fv).to(dev)
        rt, ct = torch.from_numpy(rows).to(dev), torch.from_numpy(cols).to(dev)
        fo = torch.zeros(rows.size, dtype=torch.complex128, device=dev)
        ho = torch.zeros(cols.size, dtype=torch.complex128, device=dev)
        p.ndft_rows_dev(fht.data_ptr(), rt.data_ptr(), rows.size, fo.data_ptr())
        p.ndft_cols_dev(fvt.data_ptr(), ct.data_ptr(), cols.size, ho.data_ptr())
        torch.cuda.synchronize()
        with pytest.raises(gpu.CunfftError):
            p.ndft_rows_dev(fht.data_ptr(), rt.data_ptr(), -1, fo.data_ptr())
    fo, ho = fo.cpu().numpy(), ho.cpu().numpy()
    okr = (rows >= 0) & (rows < M)
    okc = (cols >= 0) & (cols < 256)
    assert np.all(np.isnan(fo[~okr])) and np.all(np.isnan(ho[~okc]))
    ref_f = orc.ndft_trafo(N, x, fh, rows=rows[okr])
    ref_h = orc.ndft_adjoint(N, x, fv, cols=cols[okc])
    assert orc.rel_l2(ref_f, fo[okr]) <= 1e-12 and orc.rel_l2(ref_h, ho[okc]) <= 1e-12
