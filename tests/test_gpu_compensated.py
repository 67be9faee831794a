"""Compensated summation in the FP32 adjoint tiles (CUNFFT_COMPENSATED).

* Correctness: compensated FP32 plans match the oracle's NFFT (on FP32-rounded
  inputs) to 1e-5 in 2D and 3D, for each plane-tiled configuration the flag
  allows.
* Effect: with long same-sign sums -- all coefficients 1, nodes clustered so
  that thousands of them land on the cells of a few tiles -- the plain FP32
  adjoint loses accuracy with the length of the sums and the compensated one
  does not; its error against the FP64 plan is smaller.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def adjoint(gpu, N, n, m, x, fv, precision, flags=0, config=None):
    with gpu.Plan(N, x.shape[0], n=n, m=m, precision=precision,
                  flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | flags) as p:
        if config is not None:
            p.config = config
        p.x[:] = x
        p.f[:] = fv
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()
        return p.f_hat.astype(np.complex128).copy()


@pytest.mark.parametrize("case", [([64, 64], [128, 128], 4, 50000), ([32, 32, 32], [64, 64, 64], 4, 40000)])
def test_compensated_adjoint_matches_the_oracle(gpu, orc, case):
    N, n, m, M = case
    r = np.random.default_rng(31)
    d = len(N)
    x = r.uniform(-0.5, 0.5, (M, d))
    fv = r.standard_normal(M) + 1j * r.standard_normal(M)
    ref = orc.nfft_adjoint(N, n, m, x, fv.astype(np.complex64).astype(np.complex128))
    for c in range(gpu.TILE_SHAPES):
        try:
            h = adjoint(gpu, N, n, m, x, fv, "single", gpu.COMPENSATED, config=c)
        except gpu.CunfftError as e:
            print(N, "config", c, "refused:", e)
            assert e.status == 2 and c > 0
            continue
        err = orc.rel_l2(ref, h)
        print(N, "compensated config", c, "rel-l2 vs oracle", err)
        assert err <= 1e-5


@pytest.mark.parametrize("d", [2, 3])
def test_compensation_helps_long_same_sign_sums(gpu, d):
    N = [64, 64] if d == 2 else [32, 32, 32]
    n = [2 * v for v in N]
    m = 4
    M = 400000 if d == 2 else 300000
    r = np.random.default_rng(7)
    c = r.normal(0.0, 0.02, (M, d))
    x = c - np.floor(c + 0.5)
    fv = np.ones(M, dtype=np.complex128)
    h64 = adjoint(gpu, N, n, m, x, fv, "double")
    h32 = adjoint(gpu, N, n, m, x, fv, "single")
    h32c = adjoint(gpu, N, n, m, x, fv, "single", gpu.COMPENSATED)
    e_plain = np.linalg.norm(h32 - h64) / np.linalg.norm(h64)
    e_comp = np.linalg.norm(h32c - h64) / np.linalg.norm(h64)
    print(f"{d}D clustered, f = 1: FP32 plain {e_plain:.3e}, FP32 compensated {e_comp:.3e}")
    assert e_comp < e_plain
    assert e_comp <= 1e-5


def test_fp64_plans_ignore_the_flag(gpu, orc):
    N, n, m, M = [32, 32], [64, 64], 4, 5000
    r = np.random.default_rng(3)
    x = r.uniform(-0.5, 0.5, (M, 2))
    fv = r.standard_normal(M) + 1j * r.standard_normal(M)
    a = adjoint(gpu, N, n, m, x, fv, "double", gpu.COMPENSATED)
    assert orc.rel_l2(orc.nfft_adjoint(N, n, m, x, fv), a) <= 1e-12


def test_compensated_plan_without_a_fitting_configuration_is_refused(gpu):
    """FP32 3D at m = 6: no tile shape holds the doubled (Kahan) tile, so the
    plan is refused instead of running uncompensated atomics."""
    with pytest.raises(gpu.CunfftError) as e:
        gpu.Plan([16, 16, 16], 1000, m=6, precision="single",
                 flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.COMPENSATED)
    assert e.value.status == 2
    # ... and with the node-lane kernels asked for (no compensated variant)
    with pytest.raises(gpu.CunfftError) as e:
        gpu.Plan([32, 32], 1000, m=4, precision="single",
                 flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.COMPENSATED | gpu.NODE_LANES)
    assert e.value.status == 2
    # ... and with the atomic scatter asked for
    with pytest.raises(gpu.CunfftError) as e:
        gpu.Plan([32, 32], 1000, m=4, precision="single",
                 flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.COMPENSATED | gpu.ATOMIC_SCATTER)
    assert e.value.status == 2


def test_compensated_plan_refuses_the_node_lane_configuration(gpu):
    with gpu.Plan([32, 32], 1000, m=4, precision="single",
                  flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.COMPENSATED) as p:
        with pytest.raises(gpu.CunfftError) as e:
            p.config = gpu.TILE_SHAPES
        assert e.value.status == 2
        assert p.config == (0, 0)
        r = np.random.default_rng(2)
        p.x[:] = r.uniform(-0.5, 0.5, (1000, 2))
        p.copyDataToDevice()
        cfg, fwd, adj = p.autotune(reps=1)
        assert adj[gpu.TILE_SHAPES] < 0 and fwd[gpu.TILE_SHAPES] < 0
        assert gpu.TILE_SHAPES not in cfg


def test_compensated_3d_m5_matches_the_oracle(gpu, orc):
    N, n, m, M = [16, 16, 16], [32, 32, 32], 5, 20000
    r = np.random.default_rng(17)
    x = r.uniform(-0.5, 0.5, (M, 3))
    fv = r.standard_normal(M) + 1j * r.standard_normal(M)
    ref = orc.nfft_adjoint(N, n, m, x, fv.astype(np.complex64).astype(np.complex128))
    h = adjoint(gpu, N, n, m, x, fv, "single", gpu.COMPENSATED)
    err = orc.rel_l2(ref, h)
    print("3D m=5 compensated rel-l2 vs oracle", err)
    assert err <= 1e-5
