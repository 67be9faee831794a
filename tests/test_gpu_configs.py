"""Parity at the full sizes of the benchmark configurations.

cfg1  1D N=1024, n=2048, m=6, M=4096, FP64, inputs uniform[-1,1): against the
      full CPU NDFT (oracle) within the m=6 window band and the CPU NFFT
      restatement to 1e-12.
cfg2  2D radial trajectory N=512^2, n=1024^2, m=4, 512 spokes x 512 radii: node
      j = 512 s + r at r_r (cos t_s, sin t_s), t_s = s pi / 512,
      r_r = -1/2 + r/512.  The 512 nodes at r = 256 coincide at the origin, some
      with coordinate -0.0 (0 times a negative cosine), which the window rule
      u = rint(n x) - signbit(x) - m treats as negative; node 0 sits at x = -0.5
      exactly.  FP64 and FP32 against the CPU NFFT (1e-12 / 1e-5) and the full
      GPU NDFT (6.9e10 terms) within the m=4 band.
cfg3  3D N=128^3, n=256^3, m=4, M=2^21 (half uniform, half N(0,0.1^2) wrapped):
      size-independent properties -- adjointness <A fh, f> = <fh, A^H f>,
      linearity, FP32 vs FP64 <= 1e-5 -- and the FP64 NDFT on 4096 sampled
      outputs within the band.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def run_both(cu, N, n, m, x, fh, fv, precision):
    with cu.Plan(N, x.shape[0], n=n, m=m, precision=precision) as p:
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
        h = p.f_hat.astype(np.complex128).copy()
    return f, h


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


def test_cfg1_1d_against_full_ndft(gpu, orc):
    r = np.random.default_rng(101)
    N, n, m, M = [1024], [2048], 6, 4096
    x = r.uniform(-0.5, 0.5, (M, 1))
    fh = r.uniform(-1, 1, 1024) + 1j * r.uniform(-1, 1, 1024)
    fv = r.uniform(-1, 1, M) + 1j * r.uniform(-1, 1, M)
    f, h = run_both(gpu, N, n, m, x, fh, fv, "double")
    assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, fh), f) <= 1e-12
    assert orc.rel_l2(orc.nfft_adjoint(N, n, m, x, fv), h) <= 1e-12
    ef = orc.rel_l2(orc.ndft_trafo(N, x, fh), f)
    eh = orc.rel_l2(orc.ndft_adjoint(N, x, fv), h)
    print("cfg1 window error fwd", ef, "adj", eh)
    assert ef <= 5e-6 and eh <= 5e-6


def radial_nodes():
    s = np.arange(512)[:, None]
    rr = -0.5 + np.arange(512)[None, :] / 512.0
    th = s * np.pi / 512
    x0 = rr * np.cos(th)
    x1 = rr * np.sin(th)
    return np.stack([x0.ravel(), x1.ravel()], axis=1)


def test_cfg2_radial_nodes_have_the_edge_cases():
    x = radial_nodes()
    centre = (x[:, 0] == 0) & (x[:, 1] == 0)
    assert centre.sum() == 512
    assert np.signbit(x[centre]).any()          # some -0.0 coordinates
    assert x[0, 0] == -0.5 and x[0, 1] == 0.0


@pytest.mark.parametrize("precision", ["double", "single"])
def test_cfg2_radial_against_nfft_and_full_ndft(gpu, orc, precision):
    import torch

    r = np.random.default_rng(202)
    N, n, m = [512, 512], [1024, 1024], 4
    x = radial_nodes()
    M = x.shape[0]
    fh = crandn(r, 512 * 512)
    fv = crandn(r, M)
    f, h = run_both(gpu, N, n, m, x, fh, fv, precision)
    tol = 1e-12 if precision == "double" else 1e-5
    ef = orc.rel_l2(orc.nfft_trafo(N, n, m, x, q(fh, precision)), f)
    eh = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, q(fv, precision)), h)
    print("cfg2", precision, "vs oracle NFFT", ef, eh)
    assert ef <= tol and eh <= tol
    # full NDFT on the GPU (FP64 oracle kernel), same rounded inputs
    with gpu.Plan(N, M, n=n, m=m, precision="double") as p:
        p.x[:] = x
        p.copyDataToDevice()
        dev = torch.device("cuda")
        fh_t = torch.from_numpy(q(fh, precision)).to(dev)
        fv_t = torch.from_numpy(q(fv, precision)).to(dev)
        nd_f = torch.empty(M, dtype=torch.complex128, device=dev)
        nd_h = torch.empty(512 * 512, dtype=torch.complex128, device=dev)
        p.ndft_rows_dev(fh_t.data_ptr(), 0, 0, nd_f.data_ptr())
        p.ndft_cols_dev(fv_t.data_ptr(), 0, 0, nd_h.data_ptr())
        torch.cuda.synchronize()
        nd_f, nd_h = nd_f.cpu().numpy(), nd_h.cpu().numpy()
    wf, wh = orc.rel_l2(nd_f, f), orc.rel_l2(nd_h, h)
    print("cfg2", precision, "window error vs NDFT", wf, wh)
    assert wf <= 3.5e-4 and wh <= 3.5e-4


@pytest.fixture(scope="module")
def cfg3(gpu):
    r = np.random.default_rng(303)
    M = 2 ** 21
    x = np.empty((M, 3))
    x[: M // 2] = r.uniform(-0.5, 0.5, (M // 2, 3))
    c = r.normal(0.0, 0.1, (M - M // 2, 3))
    x[M // 2:] = c - np.floor(c + 0.5)
    fh = crandn(r, 128 ** 3)
    fv = crandn(r, M)
    return x, fh, fv


def test_cfg3_adjointness_linearity_and_precision(gpu, cfg3):
    x, fh, fv = cfg3
    N, n, m = [128] * 3, [256] * 3, 4
    f64, h64 = run_both(gpu, N, n, m, x, fh, fv, "double")
    # <A fh, fv> = <fh, A^H fv>
    lhs, rhs = np.vdot(fv, f64), np.vdot(h64, fh)
    print("cfg3 adjointness", abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) / abs(lhs) <= 1e-12
    # linearity of the forward: A(2 fh + i fh) = (2 + i) A fh
    f_lin, _ = run_both(gpu, N, n, m, x, (2 + 1j) * fh, fv, "double")
    assert np.linalg.norm(f_lin - (2 + 1j) * f64) / np.linalg.norm(f_lin) <= 1e-13
    f32, h32 = run_both(gpu, N, n, m, x, fh, fv, "single")
    ef = np.linalg.norm(f32 - f64) / np.linalg.norm(f64)
    eh = np.linalg.norm(h32 - h64) / np.linalg.norm(h64)
    print("cfg3 FP32 vs FP64", ef, eh)
    assert ef <= 1e-5 and eh <= 1e-5


def test_cfg3_sampled_ndft(gpu, cfg3):
    import torch

    x, fh, fv = cfg3
    N, n, m = [128] * 3, [256] * 3, 4
    f64, h64 = run_both(gpu, N, n, m, x, fh, fv, "double")
    r = np.random.default_rng(7)
    rows = np.sort(r.choice(x.shape[0], 4096, replace=False)).astype(np.int64)
    cols = np.sort(r.choice(128 ** 3, 4096, replace=False)).astype(np.int64)
    dev = torch.device("cuda")
    with gpu.Plan(N, x.shape[0], n=n, m=m, precision="double") as p:
        p.x[:] = x
        p.copyDataToDevice()
        # the inputs stay referenced until the kernels have run
        fh_t, fv_t = torch.from_numpy(fh).to(dev), torch.from_numpy(fv).to(dev)
        rows_t, cols_t = torch.from_numpy(rows).to(dev), torch.from_numpy(cols).to(dev)
        nf = torch.empty(4096, dtype=torch.complex128, device=dev)
        nh = torch.empty(4096, dtype=torch.complex128, device=dev)
        p.ndft_rows_dev(fh_t.data_ptr(), rows_t.data_ptr(), 4096, nf.data_ptr())
        p.ndft_cols_dev(fv_t.data_ptr(), cols_t.data_ptr(), 4096, nh.data_ptr())
        torch.cuda.synchronize()
    wf = np.linalg.norm(f64[rows] - nf.cpu().numpy()) / np.linalg.norm(nf.cpu().numpy())
    wh = np.linalg.norm(h64[cols] - nh.cpu().numpy()) / np.linalg.norm(nh.cpu().numpy())
    print("cfg3 window error on 4096 samples", wf, wh)
    assert wf <= 3.5e-4 and wh <= 3.5e-4


def test_cfg3_gpu_ndft_matches_the_cpu_oracle(gpu, orc, cfg3):
    """The GPU NDFT that the sampled window-error check above relies on,
    tied to the CPU oracle at the benchmark's own size: 256 rows and 256
    columns of the 2^21 x 128^3 NDFT (5.4e8 terms each way)."""
    import torch

    x, fh, fv = cfg3
    N = [128] * 3
    r = np.random.default_rng(71)
    rows = np.sort(r.choice(x.shape[0], 256, replace=False)).astype(np.int64)
    cols = np.sort(r.choice(128 ** 3, 256, replace=False)).astype(np.int64)
    dev = torch.device("cuda")
    with gpu.Plan(N, x.shape[0], n=[256] * 3, m=4, precision="double") as p:
        p.x[:] = x
        p.copyDataToDevice()
        fh_t, fv_t = torch.from_numpy(fh).to(dev), torch.from_numpy(fv).to(dev)
        rows_t, cols_t = torch.from_numpy(rows).to(dev), torch.from_numpy(cols).to(dev)
        nf = torch.empty(256, dtype=torch.complex128, device=dev)
        nh = torch.empty(256, dtype=torch.complex128, device=dev)
        p.ndft_rows_dev(fh_t.data_ptr(), rows_t.data_ptr(), 256, nf.data_ptr())
        p.ndft_cols_dev(fv_t.data_ptr(), cols_t.data_ptr(), 256, nh.data_ptr())
        torch.cuda.synchronize()
    ef = orc.rel_l2(orc.ndft_trafo(N, x, fh, rows=rows), nf.cpu().numpy())
    eh = orc.rel_l2(orc.ndft_adjoint(N, x, fv, cols=cols), nh.cpu().numpy())
    print("cfg3 GPU NDFT vs CPU oracle: rows", ef, "columns", eh)
    assert ef <= 1e-12 and eh <= 1e-12
