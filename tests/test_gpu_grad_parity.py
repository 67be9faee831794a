"""cunfft_backward_dev against the oracle in 1, 2 and 3 dimensions.

For an upstream gradient v (conjugate-Wirtinger convention), the backward of
f = A(x) fhat gives grad_fhat = A^H v (the adjoint NFFT) and
grad_x[j, t] = Re(conj(v_j) d f_j / d x_jt), the derivative of the gather of
src/cunfft_kernel.cu:514-1093 with the window index set held fixed, with the
weights recomputed (nothing stored from the forward).  The oracle's
nfft_dx / nfft_adjoint give the same quantities in FP64: FP64 plans match to
1e-12, FP32 plans to 1e-5 on FP32-rounded inputs.  Every gridding
configuration of the plan is run: for d >= 2 the plane-tiled configurations
run the binned derivative gather (k_tile_grad) on the forward's binning, the
node-lane configuration and d = 1 the node-parallel k_grad_x.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CASES = [([256], 6, 3000), ([64, 48], 4, 20000), ([24, 16, 32], 4, 20000), ([16, 16, 16], 2, 5000)]


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_backward_matches_the_oracle(gpu, orc, case, precision):
    N, m, M = CASES[case]
    d = len(N)
    n = orc.default_n(N).tolist()
    r = np.random.default_rng(400 + case)
    x = r.uniform(-0.5, 0.5, (M, d))
    fh, v = crandn(r, int(np.prod(N))), crandn(r, M)
    ct = torch.complex64 if precision == "single" else torch.complex128
    rt = torch.float32 if precision == "single" else torch.float64
    fq = fh.astype(np.complex64).astype(np.complex128) if precision == "single" else fh
    vq = v.astype(np.complex64).astype(np.complex128) if precision == "single" else v
    ref_gx = (np.conj(vq)[:, None] * orc.nfft_dx(N, n, m, x, fq)).real
    ref_gf = orc.nfft_adjoint(N, n, m, x, vq)
    tol = 1e-12 if precision == "double" else 1e-5
    dev = torch.device("cuda")
    xt = torch.from_numpy(x).to(dev)
    fht, vt = torch.from_numpy(fh).to(dev).to(ct), torch.from_numpy(v).to(dev).to(ct)
    for config in range(gpu.TUNE_CONFIGS if d >= 2 else 1):
        with gpu.Plan(N, M, n=n, m=m, precision=precision) as p:
            try:
                p.config = config
            except gpu.CunfftError as e:
                assert e.status == 2
                continue
            p.set_nodes_dev(xt.data_ptr())
            gx = torch.zeros((M, d), dtype=rt, device=dev)
            gf = torch.zeros(int(np.prod(N)), dtype=ct, device=dev)
            p.backward_dev(fht.data_ptr(), vt.data_ptr(), gf.data_ptr(), gx.data_ptr())
            torch.cuda.synchronize()
        ex = np.linalg.norm(gx.cpu().numpy().astype(np.float64) - ref_gx) / np.linalg.norm(ref_gx)
        ef = orc.rel_l2(ref_gf, gf.cpu().numpy().astype(np.complex128))
        print(N, m, precision, "config", config, "grad_x", ex, "grad_fhat", ef)
        assert ex <= tol and ef <= tol
