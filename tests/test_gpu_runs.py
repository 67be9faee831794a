"""3D adjoint paths that depend on the order of the nodes within a tile.

The binning sorts a tile's nodes by row, column unit and plane of their window
start; the spread then holds a wavefront's tile units in registers across
consecutive nodes that update the same units, and the fold sums the slots of
every chunk of a tile.  These node sets drive those paths hard:

* ``line0``: nodes on a few lines parallel to axis 0 -- every node of a tile
  shares its rows and column units, so runs are long and tiles hold several
  256-node chunks (the fold's further-chunk path);
* ``line2``: lines parallel to axis 2 -- rows fixed, column units changing;
* ``points``: many coincident nodes (identical windows) among uniform ones;
* ``cells``: nodes exactly on grid cells, x = k / n, including -0.5.

Every plane-tiled configuration (and the compensated FP32 adjoint) must give
the oracle's adjoint and forward NFFT, FP64 to 1e-12 and FP32 to 1e-5 on
FP32-rounded inputs, and two adjoint calls on one binning must agree bit for
bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, n, m, M = [16, 24, 32], [32, 48, 64], 4, 6000


def crandn(r, k):
    return r.standard_normal(k) + 1j * r.standard_normal(k)


def q(a, precision):
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


def node_set(kind, r):
    x = r.uniform(-0.5, 0.5, (M, 3))
    if kind == "line0":
        lines = r.uniform(-0.5, 0.5, (5, 3))
        x[:, 1:] = lines[r.integers(0, 5, M), 1:]
    elif kind == "line2":
        lines = r.uniform(-0.5, 0.5, (5, 3))
        x[:, :2] = lines[r.integers(0, 5, M), :2]
    elif kind == "points":
        pts = r.uniform(-0.5, 0.5, (7, 3))
        x[: M // 2] = pts[r.integers(0, 7, M // 2)]
    elif kind == "cells":
        k = r.integers(-np.array(n) // 2, np.array(n) // 2, (M, 3))
        x = k / np.array(n, dtype=np.float64)
    return x


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("kind", ["line0", "line2", "points", "cells"])
def test_sorted_runs_match_the_oracle(gpu, orc, kind, precision):
    r = np.random.default_rng({"line0": 1, "line2": 2, "points": 3, "cells": 4}[kind])
    x = node_set(kind, r)
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, M)
    ref_f = orc.nfft_trafo(N, n, m, x, q(fh, precision))
    ref_h = orc.nfft_adjoint(N, n, m, x, q(fv, precision))
    tol = 1e-12 if precision == "double" else 1e-5
    variants = [(c, 0) for c in range(3)]
    if precision == "single":
        variants += [(c, gpu.COMPENSATED) for c in range(3)]
    for config, flags in variants:
        with gpu.Plan(N, M, n=n, m=m, precision=precision,
                      flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | flags) as p:
            try:
                p.config = config
            except gpu.CunfftError as e:
                print(kind, precision, config, flags, "refused:", e)
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
            h1 = p.f_hat.copy()
            p.adjoint()
            p.copyDataToHostAd()
            eh = orc.rel_l2(ref_h, h1.astype(np.complex128))
            print(kind, precision, "config", config, "flags", flags, ef, eh)
            assert ef <= tol and eh <= tol
            assert np.array_equal(h1.view(np.uint8), p.f_hat.view(np.uint8))
