"""Inverse NFFT (cunfft_solve_dev, CGNR on the normal equations).

* A consistent overdetermined system y = A f_true (A the plan's own NFFT,
  M = 4 N_total random nodes): the solver recovers f_true, with and without
  weights, and its residual never increases (CGNR minimises |r| over growing
  Krylov spaces).
* An inconsistent system (noise added): the solution matches numpy's least
  squares on the NDFT matrix to the window error of m = 6; the matrix is
  checked against the oracle's NDFT first.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def setup(gpu, precision, N, n, m, M, seed):
    import torch

    r = np.random.default_rng(seed)
    x = r.uniform(-0.5, 0.5, (M, len(N)))
    p = gpu.Plan(N, M, n=n, m=m, precision=precision,
                 flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.NO_HOST_BUFFERS)
    xt = torch.from_numpy(x).cuda()
    p.set_nodes_dev(xt.data_ptr())
    torch.cuda.synchronize()
    return p, x, r


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("weighted", [False, True])
def test_solver_recovers_a_consistent_system(gpu, precision, weighted):
    import torch

    N, n, m = [32, 32], [64, 64], 6
    M = 4 * 32 * 32
    p, x, r = setup(gpu, precision, N, n, m, M, 41)
    cd = torch.complex128 if precision == "double" else torch.complex64
    rd = torch.float64 if precision == "double" else torch.float32
    try:
        f_true = torch.from_numpy(crandn(r, 1024)).to(cd).cuda()
        y = torch.empty(M, dtype=cd, device="cuda")
        p.transform_dev(f_true.data_ptr(), y.data_ptr())
        w = torch.from_numpy(r.uniform(0.5, 2.0, M)).to(rd).cuda() if weighted else None
        f = torch.zeros(1024, dtype=cd, device="cuda")
        torch.cuda.synchronize()
        tol = 1e-12 if precision == "double" else 1e-6
        iters, hist = p.solve_dev(y.data_ptr(), f.data_ptr(), w.data_ptr() if weighted else 0,
                                  max_iter=300, tol=tol)
        err = float((f - f_true).abs().norm() / f_true.abs().norm())
        print(precision, "weighted" if weighted else "plain", "iterations", iters,
              "final residual", hist[-1], "solution error", err)
        assert iters < 300 and hist[-1] <= tol
        assert err <= (1e-8 if precision == "double" else 2e-4)
        h = np.array(hist)
        assert np.all(h[1:] <= h[:-1] * (1 + 1e-6) + (1e-12 if precision == "double" else 1e-6))
    finally:
        p.finalize()


def test_solver_matches_least_squares_on_the_ndft_matrix(gpu, orc):
    import torch

    N, n, m = [12, 16], [24, 32], 6
    M = 600
    p, x, r = setup(gpu, "double", N, n, m, M, 43)
    try:
        k0 = np.arange(-N[0] // 2, N[0] // 2)
        k1 = np.arange(-N[1] // 2, N[1] // 2)
        K0, K1 = np.meshgrid(k0, k1, indexing="ij")
        A = np.exp(-2j * np.pi * (np.outer(x[:, 0], K0.ravel()) + np.outer(x[:, 1], K1.ravel())))
        probe = crandn(r, A.shape[1])
        assert orc.rel_l2(orc.ndft_trafo(N, x, probe), A @ probe) <= 1e-12   # the matrix is the NDFT
        y = A @ crandn(r, A.shape[1]) + 0.3 * crandn(r, M)                  # inconsistent
        ref = np.linalg.lstsq(A, y, rcond=None)[0]
        yt = torch.from_numpy(y).cuda()
        f = torch.zeros(A.shape[1], dtype=torch.complex128, device="cuda")
        iters, hist = p.solve_dev(yt.data_ptr(), f.data_ptr(), 0, max_iter=500, tol=1e-14)
        err = np.linalg.norm(f.cpu().numpy() - ref) / np.linalg.norm(ref)
        print("least squares: iterations", iters, "residual", hist[-1], "vs lstsq", err)
        assert err <= 1e-5
    finally:
        p.finalize()


def test_solver_arguments(gpu):
    import torch

    p, x, r = setup(gpu, "double", [16, 16], [32, 32], 4, 300, 44)
    try:
        y = torch.zeros(300, dtype=torch.complex128, device="cuda")
        f = torch.zeros(256, dtype=torch.complex128, device="cuda")
        iters, hist = p.solve_dev(y.data_ptr(), f.data_ptr(), 0, max_iter=5, tol=1e-10)
        assert iters == 0 and float(f.abs().max()) == 0.0     # y = 0: f = 0 is exact
        with pytest.raises(gpu.CunfftError):
            p.solve_dev(0, f.data_ptr())
    finally:
        p.finalize()
