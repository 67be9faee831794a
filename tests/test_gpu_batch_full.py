"""The batched workload at its full size (bench.py's batch, rebuilt from the
same seeds): 96 uneven 2D problems N_b^2, N_b in {64, 96, 128, 192, 256},
sigma = 2 exactly (n_b = 2 N_b: 192 and 384 for N_b = 96 and 192, on the
radix-3 passes), m = 4, M_b in [4096, 65536], 8 right-hand sides, FP32, one
graph launch per direction.

* every (problem, right-hand side) of the forward and the adjoint against
  FP64 single plans on the same (FP32) inputs, <= 1e-5 relative l2, for every
  batch gridding configuration;
* three problems -- N_b = 256 with the most nodes, one of N_b in {96, 192},
  one of N_b = 64 -- against the CPU oracle's NFFT, <= 1e-5.
"""
import os
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.fixture(scope="module")
def workload(gpu):
    from cunfft_amd.batch import Batch

    N, n, M, problem = bench.batch_workload()
    dev = torch.device("cuda")
    b = Batch(N, M, m=bench.BATCH_M, n_rhs=bench.BATCH_R, n=n, precision="single")
    xs = bench.load_batch(b, problem, dev)
    # FP64 references: single FP64 plans on the batch's own (FP32) inputs
    refs = []
    for i in range(b.P):
        fi, _, ai, _ = b.tensors(i)
        with gpu.Plan(list(N[i]), int(M[i]), n=list(n[i]), m=bench.BATCH_M, precision="double",
                      flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.NO_HOST_BUFFERS) as p:
            p.set_nodes_dev(xs[i].data_ptr())
            fo = torch.empty((b.n_rhs, int(M[i])), dtype=torch.complex128, device=dev)
            ao = torch.empty((b.n_rhs, int(N[i][0] * N[i][1])), dtype=torch.complex128, device=dev)
            src_f = fi.to(torch.complex128).contiguous()
            src_a = ai.to(torch.complex128).contiguous()
            for k in range(b.n_rhs):
                p.transform_dev(src_f[k].data_ptr(), fo[k].data_ptr())
                p.adjoint_dev(src_a[k].data_ptr(), ao[k].data_ptr())
            torch.cuda.synchronize()
        refs.append((fo, ao))
    yield b, N, n, M, problem, xs, refs
    b.close()


def rel(a, b):
    return float(((a.to(torch.complex128) - b).norm() / b.norm()).item())


@pytest.mark.parametrize("config", [0, 1, 2, 3, "autotuned"])
def test_every_problem_and_rhs_matches_fp64_plans(workload, config):
    b, N, n, M, problem, xs, refs = workload
    if config == "autotuned":
        cfg, _, _ = b.autotune(reps=2)
    else:
        try:
            b.config = config
        except Exception as e:   # a tile shape that does not fit some problem
            assert getattr(e, "status", None) == 2, e
            print("batch config", config, "not available:", e)
            return
        cfg = b.config
    b.execute()
    torch.cuda.synchronize()
    worst = 0.0
    for i in range(b.P):
        _, fo, _, ao = b.tensors(i)
        rf, ra = refs[i]
        for k in range(b.n_rhs):
            e = max(rel(fo[k], rf[k]), rel(ao[k], ra[k]))
            worst = max(worst, e)
            assert e <= 1e-5, (i, k, list(N[i]), int(M[i]), e)
    print("batch config", cfg, "worst FP32 vs FP64 plan rel-l2 over 96 x 8", worst)


def test_three_problems_match_the_oracle(workload, orc):
    b, N, n, M, problem, xs, refs = workload
    b.config = 0
    b.execute()
    torch.cuda.synchronize()
    sizes = N[:, 0]
    big = [i for i in range(b.P) if sizes[i] == 256]
    picks = [max(big, key=lambda i: M[i]),
             next(i for i in range(b.P) if sizes[i] in (96, 192)),
             next(i for i in range(b.P) if sizes[i] == 64)]
    for i in picks:
        x, fh, fv = problem(i)
        _, fo, _, ao = b.tensors(i)
        Ni, ni = list(N[i]), list(n[i])
        for k in (0, b.n_rhs - 1):
            ref_f = orc.nfft_trafo(Ni, ni, bench.BATCH_M, x, fh[k].astype(np.complex128))
            ref_h = orc.nfft_adjoint(Ni, ni, bench.BATCH_M, x, fv[k].astype(np.complex128))
            ef = orc.rel_l2(ref_f, fo[k].cpu().numpy())
            eh = orc.rel_l2(ref_h, ao[k].cpu().numpy())
            print("problem", i, Ni, "n", ni, "M", int(M[i]), "rhs", k, "vs oracle", ef, eh)
            assert ef <= 1e-5 and eh <= 1e-5
