"""Gradients of the NFFT layer (cunfft_amd.autograd over cunfft_backward_dev).

Workload (the differentiation study): 2D, N = 256^2, n = 512^2, m = 4,
M = 65536 nodes uniform in [-0.5, 0.5)^2, loss L = 0.5 ||A(x) fhat - y||^2 with
y = A(x + 0.01 N(0,1), wrapped) fhat_true and fhat, fhat_true re/im N(0,1).

Checks:
* dL/dx (derivative-of-Gaussian gridding) against FP64 central differences
  with step 1e-6 on 64 seeded coordinates, <= 1e-6 relative (rel-l2 over the
  64 entries), skipping coordinates within 10 steps of a window jump (the
  window start u = rint(n x) - signbit(x) - m changes at half-integer n x and
  at x = 0, where the truncated window makes L non-differentiable);
* dL/dfhat (adjoint NFFT) against central differences on 32 seeded
  coefficients (real and imaginary parts; L is quadratic in fhat, so the
  difference is exact up to roundoff and a step of 1e-3 keeps roundoff small);
* dL/dx against the analytic gradient of the NDFT model,
  Re(conj(r_j) sum_k fhat_k (-2 pi i k_t) e^{-2 pi i k.x_j}), within the
  m = 4 window band (5e-4);
* FP32 gradients against FP64 ones (precision error, rel-l2 <= 1e-4).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = [256, 256]
NN = [512, 512]
M_CUT = 4
M = 65536


@pytest.fixture(scope="module")
def problem(gpu):
    import torch

    from cunfft_amd.autograd import NFFTOperator

    r = np.random.default_rng(2024)
    x = r.uniform(-0.5, 0.5, (M, 2))
    xn = x + 0.01 * r.standard_normal((M, 2))
    xn = xn - np.floor(xn + 0.5)
    Nt = int(np.prod(N))
    fhat = r.standard_normal(Nt) + 1j * r.standard_normal(Nt)
    fhat_true = r.standard_normal(Nt) + 1j * r.standard_normal(Nt)
    op64 = NFFTOperator(N, M, n=NN, m=M_CUT, precision="double")
    op32 = NFFTOperator(N, M, n=NN, m=M_CUT, precision="single")
    dev = torch.device("cuda")
    y = op64.forward(torch.from_numpy(fhat_true).to(dev), torch.from_numpy(xn).to(dev))
    torch.cuda.synchronize()
    yield {"x": x, "fhat": fhat, "y": y, "op64": op64, "op32": op32, "dev": dev, "rng": r}
    op64.close()
    op32.close()


def loss_and_grads(p, precision):
    import torch

    from cunfft_amd.autograd import nfft

    op = p["op64"] if precision == "double" else p["op32"]
    cd = torch.complex128 if precision == "double" else torch.complex64
    fh = torch.from_numpy(p["fhat"]).to(p["dev"]).to(cd).requires_grad_(True)
    x = torch.from_numpy(p["x"]).to(p["dev"]).requires_grad_(True)
    f = nfft(fh, x, op)
    L = 0.5 * ((f.to(torch.complex128) - p["y"]).abs() ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return float(L), fh.grad.detach().to(torch.complex128).cpu().numpy(), x.grad.detach().cpu().numpy()


def loss_at(p, fhat, x):
    import torch

    f = p["op64"].forward(torch.from_numpy(fhat).to(p["dev"]), torch.from_numpy(x).to(p["dev"]))
    return float(0.5 * ((f - p["y"]).abs() ** 2).sum())


def smooth_coords(x, n, h, count, rng):
    """Seeded (node, axis) pairs whose window start is constant on [x - 10h, x + 10h]."""
    out = []
    while len(out) < count:
        j, t = int(rng.integers(M)), int(rng.integers(2))
        v = x[j, t]
        frac = n[t] * v - np.floor(n[t] * v)
        if abs(frac - 0.5) > 10 * n[t] * h and abs(v) > 10 * h and (j, t) not in out:
            out.append((j, t))
    return out


def test_grad_x_matches_central_differences_fp64(problem):
    p = problem
    _, _, gx = loss_and_grads(p, "double")
    h = 1e-6
    coords = smooth_coords(p["x"], NN, h, 64, np.random.default_rng(5))
    fd, ad = [], []
    for j, t in coords:
        xp = p["x"].copy()
        xm = p["x"].copy()
        xp[j, t] += h
        xm[j, t] -= h
        fd.append((loss_at(p, p["fhat"], xp) - loss_at(p, p["fhat"], xm)) / (2 * h))
        ad.append(gx[j, t])
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(fd - ad) / np.linalg.norm(ad)
    print("dL/dx: FD vs autograd rel-l2", rel, "max |g|", np.abs(ad).max())
    assert rel <= 1e-6


def test_grad_fhat_matches_central_differences_fp64(problem):
    p = problem
    _, gf, _ = loss_and_grads(p, "double")
    h = 1e-3
    rng = np.random.default_rng(6)
    idx = rng.choice(len(p["fhat"]), 32, replace=False)
    fd, ad = [], []
    for k in idx:
        for unit in (1.0, 1j):
            fp = p["fhat"].copy()
            fm = p["fhat"].copy()
            fp[k] += h * unit
            fm[k] -= h * unit
            fd.append((loss_at(p, fp, p["x"]) - loss_at(p, fm, p["x"])) / (2 * h))
            ad.append(gf[k].real if unit == 1.0 else gf[k].imag)
    fd, ad = np.array(fd), np.array(ad)
    rel = np.linalg.norm(fd - ad) / np.linalg.norm(ad)
    print("dL/dfhat: FD vs autograd rel-l2", rel)
    assert rel <= 1e-6


def test_grad_x_matches_ndft_gradient(problem):
    """Analytic gradient of the NDFT model on 64 nodes (GPU FP64 NDFT rows)."""
    import torch

    p = problem
    _, _, gx = loss_and_grads(p, "double")
    rng = np.random.default_rng(7)
    rows = np.sort(rng.choice(M, 64, replace=False)).astype(np.int64)
    op = p["op64"]
    dev = p["dev"]
    op.set_nodes(torch.from_numpy(p["x"]).to(dev))
    rows_t = torch.from_numpy(rows).to(dev)
    k = [np.arange(-N[0] // 2, N[0] // 2), np.arange(-N[1] // 2, N[1] // 2)]
    K0, K1 = np.meshgrid(k[0], k[1], indexing="ij")
    out = {}
    for name, coef in (("f", p["fhat"]), ("d0", p["fhat"] * (-2j * np.pi) * K0.ravel()),
                       ("d1", p["fhat"] * (-2j * np.pi) * K1.ravel())):
        res = torch.empty(64, dtype=torch.complex128, device=dev)
        c = torch.from_numpy(np.ascontiguousarray(coef)).to(dev)
        op.plan.ndft_rows_dev(c.data_ptr(), rows_t.data_ptr(), 64, res.data_ptr(),
                              torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        out[name] = res.cpu().numpy()
    r = out["f"] - p["y"].cpu().numpy()[rows]
    g_ndft = np.stack([(np.conj(r) * out["d0"]).real, (np.conj(r) * out["d1"]).real], axis=1)
    rel = np.linalg.norm(gx[rows] - g_ndft) / np.linalg.norm(g_ndft)
    print("dL/dx: NFFT autograd vs analytic NDFT gradient rel-l2", rel)
    assert rel <= 5e-4


def test_fp32_gradients_match_fp64(problem):
    p = problem
    L64, gf64, gx64 = loss_and_grads(p, "double")
    L32, gf32, gx32 = loss_and_grads(p, "single")
    rf = np.linalg.norm(gf32 - gf64) / np.linalg.norm(gf64)
    rx = np.linalg.norm(gx32 - gx64) / np.linalg.norm(gx64)
    print("FP32 vs FP64: loss", abs(L32 - L64) / L64, "dL/dfhat", rf, "dL/dx", rx)
    assert abs(L32 - L64) / L64 <= 1e-5
    assert rf <= 1e-4 and rx <= 1e-4


def test_backward_entry_point_outputs_are_optional(problem):
    """cunfft_backward_dev with one output NULL computes the other one alone."""
    import torch

    p = problem
    op = p["op64"]
    dev = p["dev"]
    fh = torch.from_numpy(p["fhat"]).to(dev)
    x = torch.from_numpy(p["x"]).to(dev)
    f = op.forward(fh, x)
    gfull = op.backward(fh, x, f - p["y"])
    gonly_f = op.backward(fh, x, f - p["y"], need_fhat=True, need_x=False)
    gonly_x = op.backward(fh, x, f - p["y"], need_fhat=False, need_x=True)
    torch.cuda.synchronize()
    assert gonly_f[1] is None and gonly_x[0] is None
    # each call re-sets (and re-bins) the nodes; the binning is a stable sort,
    # so the adjoint sums run in the same order and agree bit for bit
    assert torch.equal(gfull[0], gonly_f[0])
    assert torch.equal(gfull[1], gonly_x[1])
