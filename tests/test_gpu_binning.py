"""The node binning is a stable device sort: bit-identical adjoints.

The adjoint sums the nodes of each grid tile in sorted order.  The binning is
a stable LSD radix sort of (tile, window start within the tile) over the node
indices, so the order -- and with it every bit of the plane-tiled adjoint --
depends on the nodes only: two binnings of the same nodes (a second
set_nodes_dev, a second plan, or another stream) give bit-identical results.
The reference's adjoint adds with CAS-loop atomics in scheduling order
(src/cunfft_kernel.cu:97-130, 1098-1197); its bits vary run to run.

The node-lane configuration (CUNFFT_TILE_SHAPES) accumulates with LDS and
global float atomics and is exempt; it is checked against the oracle only.

Node sets: the radial trajectory of the 2D benchmark (512 coincident nodes at
the centre, half of them with -0.0 coordinates), coincident points among
uniform nodes in 3D, and the clustered half of the 3D benchmark.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def radial(S, R):
    th = np.arange(S) * np.pi / S
    rr = -0.5 + np.arange(R) / R
    x = np.empty((S * R, 2))
    x[:, 0] = (rr[None, :] * np.cos(th)[:, None]).reshape(-1)
    x[:, 1] = (rr[None, :] * np.sin(th)[:, None]).reshape(-1)
    return x


def points3(r, M):
    x = r.uniform(-0.5, 0.5, (M, 3))
    pts = r.uniform(-0.5, 0.5, (7, 3))
    x[: M // 2] = pts[r.integers(0, 7, M // 2)]
    return x


def clustered3(r, M):
    x = r.uniform(-0.5, 0.5, (M, 3))
    c = r.normal(0.0, 0.1, (M // 2, 3))
    x[M // 2:] = c - np.floor(c + 0.5)
    return x


CASES = {
    "radial2d": ([128, 128], 4, lambda r: radial(256, 256)),
    "points3d": ([16, 24, 32], 4, lambda r: points3(r, 6000)),
    "clustered3d": ([32, 32, 32], 4, lambda r: clustered3(r, 60000)),
}


def adjoint_twice(gpu, N, m, x, fv, precision, config):
    """Adjoint of the same nodes binned three times: twice in one plan (the
    nodes set again in between, on another stream), once in a second plan."""
    dev = torch.device("cuda")
    ct = torch.complex64 if precision == "single" else torch.complex128
    xt = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    ft = torch.from_numpy(fv).to(dev).to(ct)
    out = []
    with gpu.Plan(N, x.shape[0], m=m, precision=precision) as p, \
            gpu.Plan(N, x.shape[0], m=m, precision=precision) as q:
        for plan in (p, q):
            try:
                plan.config = config
            except gpu.CunfftError as e:
                assert e.status == 2
                return None
        side = torch.cuda.Stream()
        for plan, stream in ((p, None), (p, side), (q, None)):
            h = torch.zeros(int(np.prod(N)), dtype=ct, device=dev)
            s = stream.cuda_stream if stream is not None else None
            torch.cuda.synchronize()
            plan.set_nodes_dev(xt.data_ptr(), s)
            plan.adjoint_dev(ft.data_ptr(), h.data_ptr(), s)
            torch.cuda.synchronize()
            out.append(h.cpu())
    return out


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_rebinning_gives_bit_identical_adjoints(gpu, orc, case, precision):
    N, m, make = CASES[case]
    r = np.random.default_rng(len(case))
    x = make(r)
    M = x.shape[0]
    fv = r.standard_normal(M) + 1j * r.standard_normal(M)
    n = orc.default_n(N).tolist()
    fq = fv.astype(np.complex64).astype(np.complex128) if precision == "single" else fv
    ref = orc.nfft_adjoint(N, n, m, x, fq)
    tol = 1e-12 if precision == "double" else 1e-5
    for config in range(gpu.TILE_SHAPES):
        res = adjoint_twice(gpu, N, m, x, fv, precision, config)
        if res is None:
            continue
        h1, h2, h3 = res
        err = orc.rel_l2(ref, h1.numpy().astype(np.complex128))
        print(case, precision, "config", config, "rel-l2 vs oracle", err)
        assert err <= tol
        assert torch.equal(h1, h2), "re-binning in the same plan changed the adjoint bits"
        assert torch.equal(h1, h3), "a second plan's binning changed the adjoint bits"


def test_node_lane_configuration_matches_the_oracle(gpu, orc):
    N, m, make = CASES["radial2d"]
    x = make(None)
    r = np.random.default_rng(9)
    fv = r.standard_normal(x.shape[0]) + 1j * r.standard_normal(x.shape[0])
    res = adjoint_twice(gpu, N, m, x, fv, "double", gpu.TILE_SHAPES)
    ref = orc.nfft_adjoint(N, orc.default_n(N).tolist(), m, x, fv)
    for h in res:
        assert orc.rel_l2(ref, h.numpy()) <= 1e-12
