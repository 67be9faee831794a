"""Gridding kernel configurations and the autotuner (cunfft_autotune,
cunfft_set_config).

Every configuration -- the three tile shapes of the plane-tiled kernels and
the node-lane kernels -- must give the oracle's NFFT (FP64 to 1e-12, FP32 to
the FP32 bound on FP32-rounded inputs), and the autotuner must keep the
configuration it timed fastest per direction -- or, when the two directions'
fastest differ, one of them on both directions (a step sharing one binning)
-- leaving a plan that still transforms correctly.  Node sets: uniform, and clustered (N(0, 0.05^2) wrapped), which
piles many chunks onto a few tiles.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def nodes(r, M, d, kind):
    if kind == "uniform":
        return r.uniform(-0.5, 0.5, (M, d))
    c = r.normal(0.0, 0.05, (M, d))
    return c - np.floor(c + 0.5)


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


def fwd_adj(p, x, fh, fv):
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


CASES = [([64, 64], [128, 128], 4, 40000), ([24, 32, 16], [64, 64, 32], 4, 30000),
         ([32, 32, 32], [64, 64, 64], 3, 20000)]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("kind", ["uniform", "clustered"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_every_configuration_matches_the_oracle(gpu, orc, precision, kind, case):
    N, n, m, M = CASES[case]
    r = np.random.default_rng(40 + case)
    x = nodes(r, M, len(N), kind)
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, M)
    ref_f = orc.nfft_trafo(N, n, m, x, q(fh, precision))
    ref_h = orc.nfft_adjoint(N, n, m, x, q(fv, precision))
    tol = 1e-12 if precision == "double" else 1e-5
    with gpu.Plan(N, M, n=n, m=m, precision=precision) as p:
        assert p.config == (0, 0)
        for c in range(gpu.TUNE_CONFIGS):
            try:
                p.config = c
            except gpu.CunfftError as e:
                # a tile shape whose LDS tile does not fit is refused, nothing else;
                # the plan keeps its configuration and stays usable
                print(N, precision, "config", c, "refused:", e)
                assert e.status == 2 and 0 < c < gpu.TILE_SHAPES
                assert p.config == (c - 1, c - 1)
                continue
            assert p.config == (c, c)
            f, h = fwd_adj(p, x, fh, fv)
            ef, eh = orc.rel_l2(ref_f, f), orc.rel_l2(ref_h, h)
            print(N, kind, precision, "config", c, ef, eh)
            assert ef <= tol and eh <= tol, (c, ef, eh)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_autotune_keeps_the_fastest_configuration(gpu, orc, precision):
    import torch

    N, n, m, M = [64, 64, 64], [128, 128, 128], 4, 200000
    r = np.random.default_rng(77)
    x = nodes(r, M, 3, "uniform")
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, M)
    with gpu.Plan(N, M, n=n, m=m, precision=precision) as p:
        p.x[:] = x
        p.copyDataToDevice()
        chosen, ms_f, ms_a = p.autotune(reps=3, stream=torch.cuda.current_stream().cuda_stream)
        print(precision, "autotune", chosen, ms_f, ms_a)
        assert len(ms_f) == len(ms_a) == gpu.TUNE_CONFIGS
        fastest = []
        for d, ms in enumerate((ms_f, ms_a)):
            valid = [c for c in range(gpu.TUNE_CONFIGS) if ms[c] >= 0]
            assert 0 in valid and gpu.TILE_SHAPES in valid
            assert chosen[d] in valid
            fastest.append(min(valid, key=lambda c: ms[c]))
        b0, b1 = fastest
        assert tuple(chosen) in {(b0, b1), (b1, b1), (b0, b0)}
        assert p.config == chosen
        f, h = fwd_adj(p, x, fh, fv)
    tol = 1e-12 if precision == "double" else 1e-5
    assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, q(fh, precision)), f) <= tol
    assert orc.rel_l2(orc.nfft_adjoint(N, n, m, x, q(fv, precision)), h) <= tol


@pytest.mark.parametrize("precision", ["double", "single"])
def test_forward_and_adjoint_configurations_can_differ(gpu, orc, precision):
    """Each direction keeps its own binning when the configurations differ;
    re-setting the nodes rebuilds both."""
    N, n, m, M = [32, 32, 32], [64, 64, 64], 4, 30000
    r = np.random.default_rng(78)
    tol = 1e-12 if precision == "double" else 1e-5
    with gpu.Plan(N, M, n=n, m=m, precision=precision) as p:
        for cf, ca in ((1, gpu.TILE_SHAPES), (gpu.TILE_SHAPES, 0), (0, 1)):
            p.config = (cf, ca)
            assert p.config == (cf, ca)
            for kind in ("uniform", "clustered"):
                x = nodes(r, M, 3, kind)
                fh = crandn(r, int(np.prod(N)))
                fv = crandn(r, M)
                f, h = fwd_adj(p, x, fh, fv)
                ef = orc.rel_l2(orc.nfft_trafo(N, n, m, x, q(fh, precision)), f)
                eh = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, q(fv, precision)), h)
                print(precision, (cf, ca), kind, ef, eh)
                assert ef <= tol and eh <= tol


def test_one_dimensional_plans_have_one_configuration(gpu):
    with gpu.Plan([256], 1000, n=[512], m=4) as p:
        p.x[:] = np.random.default_rng(1).uniform(-0.5, 0.5, (1000, 1))
        p.copyDataToDevice()
        chosen, ms_f, ms_a = p.autotune(reps=2)
        assert chosen == (0, 0)
        for ms in (ms_f, ms_a):
            assert ms[0] >= 0 and all(v == -1 for v in ms[1:])
        with pytest.raises(gpu.CunfftError):
            p.config = 1
        assert p.config == (0, 0)
        with pytest.raises(gpu.CunfftError):
            p.autotune(directions=4)
