#!/usr/bin/env python3
"""Forward + adjoint NFFT throughput on one MI355X.

Workload: 3D NFFT, N = 128^3, n = 256^3, m = 4, M = 2^21 nodes (half uniform in
[-0.5,0.5)^3, half N(0, 0.1^2) per axis wrapped onto the torus), fhat and f
re/im N(0,1).  One step = one forward NFFT plus one adjoint NFFT on inputs
resident in HBM (nodes binned once, before timing; x is fixed across steps).
The headline runs in FP32 complex with the FP64 index/fraction split and FP64
roll-off factors; the same steps are timed in full FP64 beside it.

Beside it (the "batched_2d" object): 96 uneven 2D FP32 problems, N_b x N_b
with N_b in {64, 96, 128, 192, 256}, sigma = 2 (n_b = 2 N_b), m = 4, M_b
uniform in [4096, 65536] nodes, 8 right-hand sides each; one step = forward +
adjoint of every problem and right-hand side, as one graph launch of the
batched kernels, against a loop of single-plan launches over the same data.
Both are autotuned before timing.

Prints one JSON line (the last line of stdout).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N3 = [128, 128, 128]
NN3 = [256, 256, 256]
M_NODES = 2 ** 21
CUT = 4
HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix-core peak
MFMA_FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64 matrix-core peak


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(seed=20240603):
    r = np.random.default_rng(seed)
    half = M_NODES // 2
    x = np.empty((M_NODES, 3))
    x[:half] = r.uniform(-0.5, 0.5, (half, 3))
    c = r.normal(0.0, 0.1, (M_NODES - half, 3))
    x[half:] = c - np.floor(c + 0.5)          # wrap onto [-0.5, 0.5)
    Nt = int(np.prod(N3))
    fh = r.standard_normal(Nt) + 1j * r.standard_normal(Nt)
    fv = r.standard_normal(M_NODES) + 1j * r.standard_normal(M_NODES)
    return x, fh, fv


def grid_bytes(cbytes):
    """Algorithmic bytes of one gridding launch: the grid once plus the nodes."""
    n_total = int(np.prod(NN3))
    return cbytes * n_total + M_NODES * (3 * 8 + cbytes)


def time_plan(cu, precision, x, fh, fv, steps, warmup):
    import torch

    p = cu.Plan(N3, M_NODES, n=NN3, m=CUT, precision=precision)
    p.x[:] = x
    p.f_hat[:] = fh
    p.f[:] = fv
    p.copyDataToDevice()
    p.copyDataToDeviceAd()
    dp = p.device_pointers
    # gridding kernels chosen by the autotuner on these nodes (untimed, plan time)
    config, ms_f, ms_a = p.autotune(reps=3)
    torch.cuda.synchronize()
    p.copyDataToDevice()
    p.copyDataToDeviceAd()
    # warm-up (also builds the node binning once)
    for _ in range(warmup):
        p.transform_dev(dp["f_hat"], dp["f"])
        p.adjoint_dev(dp["f"], dp["f_hat"])
    torch.cuda.synchronize()
    p.copyDataToDevice()
    p.copyDataToDeviceAd()
    torch.cuda.synchronize()
    # timed region: K steps of forward + adjoint, device-resident data, binning warm
    t0 = time.perf_counter()
    for _ in range(steps):
        p.transform_dev(dp["f_hat"], dp["f"])
        p.adjoint_dev(dp["f"], dp["f_hat"])
    torch.cuda.synchronize()  # device-wide: waits for the plan's own stream too
    ms_per_step = 1e3 * (time.perf_counter() - t0) / steps
    # cold binning: the nodes are re-set (and re-sorted) before every step
    cold_steps = max(2, steps // 4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(cold_steps):
        p.set_nodes_dev(dp["x"])
        p.transform_dev(dp["f_hat"], dp["f"])
        p.adjoint_dev(dp["f"], dp["f_hat"])
    torch.cuda.synchronize()
    cold_ms = 1e3 * (time.perf_counter() - t0) / cold_steps
    # per-kernel device times with HIP events on the plan stream
    p._p.flags = p._p.flags | cu.MEASURE_TIME
    rec = {k: [] for k in ("fwd_fft", "fwd_gather", "adj_grid", "adj_fold", "adj_fft")}
    for _ in range(max(3, min(steps, 10))):
        p.transform()
        t = p.times
        rec["fwd_fft"].append(t["time_FFT"])
        rec["fwd_gather"].append(t["time_CONV"])
        p.adjoint()
        t = p.times
        rec["adj_grid"].append(t["time_CONV"])
        rec["adj_fold"].append(t["time_else"])
        rec["adj_fft"].append(t["time_FFT"])
    p._p.flags = p._p.flags & ~cu.MEASURE_TIME
    med = {k: 1e3 * float(np.median(v)) for k, v in rec.items()}
    bwd = time_backward(p, precision, steps)
    return p, {"config": list(config), "autotune_fwd_ms": ms_f, "autotune_adj_ms": ms_a,
               "ms_per_step": ms_per_step, "cold_binning_ms_per_step": cold_ms,
               "fwd_fft_ms": med["fwd_fft"], "fwd_gather_ms": med["fwd_gather"],
               "adj_scatter_ms": med["adj_grid"] - med["adj_fold"], "adj_fold_ms": med["adj_fold"],
               "adj_fft_ms": med["adj_fft"], **bwd}


def time_backward(p, precision, steps):
    """Backward of f = A(x) fhat on the same plan (binning warm): the full
    backward (grad_x from the forward grid of fhat -- roll-off, FFT, derivative
    gather -- then grad_fhat = A^H grad_f, one adjoint NFFT) and the
    x-gradient alone, with grad_f = the plan's f."""
    import torch

    dp = p.device_pointers
    real = torch.float32 if precision == "single" else torch.float64
    cplx = torch.complex64 if precision == "single" else torch.complex128
    gx = torch.empty(M_NODES * 3, dtype=real, device="cuda")
    gfh = torch.empty(int(np.prod(N3)), dtype=cplx, device="cuda")
    reps = max(3, min(steps, 10))
    out = {}
    for key, call in (("backward_ms", lambda: p.backward_dev(dp["f_hat"], dp["f"], gfh.data_ptr(),
                                                             gx.data_ptr())),
                      ("grad_x_ms", lambda: p.grad_x_dev(dp["f_hat"], dp["f"], gx.data_ptr()))):
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        torch.cuda.synchronize()
        out[key] = 1e3 * (time.perf_counter() - t0) / reps
    return out


def pcie_inclusive(cu, x, fh, fv, steps=4):
    """FP32 steps through the host-buffer boundary (copy in, transform, copy out,
    then the same for the adjoint), the reference's call sequence: synchronous,
    and with CUNFFT_ASYNC_COPY on two plans whose copies and kernels overlap."""
    def make(flags):
        p = cu.Plan(N3, M_NODES, n=NN3, m=CUT, precision="single", flags=flags)
        p.x[:] = x
        p.f_hat[:] = fh
        p.f[:] = fv
        return p

    def step(p):
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()

    base = cu.CUFFT_INIT | cu.CUFFT_BOTH
    p = make(base)
    step(p)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(p)
    ms_sync = 1e3 * (time.perf_counter() - t0) / steps
    p.finalize()
    pa, pb = make(base | cu.ASYNC_COPY), make(base | cu.ASYNC_COPY)

    def two():
        # forward of both plans queued before either result is read back
        for q in (pa, pb):
            q.copyDataToDevice()
            q.transform()
        for q in (pa, pb):
            q.copyDataToHost()
        for q in (pa, pb):
            q.copyDataToDeviceAd()
            q.adjoint()
        for q in (pa, pb):
            q.copyDataToHostAd()

    two()
    t0 = time.perf_counter()
    for _ in range(steps):
        two()
    ms_async = 1e3 * (time.perf_counter() - t0) / (2 * steps)
    pa.finalize()
    pb.finalize()
    return {"sync_ms_per_step": ms_sync, "sync_nodes_per_s": M_NODES / (ms_sync * 1e-3),
            "async_two_plans_ms_per_step": ms_async,
            "async_two_plans_nodes_per_s": M_NODES / (ms_async * 1e-3),
            # x twice (d doubles per node), f twice and f_hat twice (FP32 complex)
            "bytes_over_pcie_per_step": 2 * 24 * M_NODES + 2 * 8 * M_NODES
            + 2 * 8 * int(np.prod(N3))}


def accuracy(cu, p32, p64, x, fh, n_sample=4096):
    """Precision and window error on the bench inputs (GPU, untimed)."""
    import torch

    out = {}
    r = np.random.default_rng(7)
    rows = np.sort(r.choice(M_NODES, n_sample, replace=False)).astype(np.int64)
    for p in (p32, p64):
        p.copyDataToDevice()
        p.transform()
    f32 = p32.copy_f_ToHost().astype(np.complex128)
    f64 = p64.copy_f_ToHost()
    out["fwd_fp32_vs_fp64_rel_l2"] = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))
    rows_t = torch.from_numpy(rows).cuda()
    nd = torch.empty(n_sample, dtype=torch.complex128, device="cuda")
    p64.ndft_rows_dev(p64.device_pointers["f_hat"], rows_t.data_ptr(), n_sample, nd.data_ptr())
    torch.cuda.synchronize()
    nd = nd.cpu().numpy()
    out["fwd_fp64_vs_ndft_rel_l2_sample"] = float(np.linalg.norm(f64[rows] - nd) / np.linalg.norm(nd))
    out["fwd_fp32_vs_ndft_rel_l2_sample"] = float(np.linalg.norm(f32[rows] - nd) / np.linalg.norm(nd))
    for p in (p32, p64):
        p.copyDataToDeviceAd()
        p.adjoint()
    h32 = p32.copy_f_hat_ToHost().astype(np.complex128)
    h64 = p64.copy_f_hat_ToHost()
    out["adj_fp32_vs_fp64_rel_l2"] = float(np.linalg.norm(h32 - h64) / np.linalg.norm(h64))
    cols = np.sort(r.choice(int(np.prod(N3)), n_sample, replace=False)).astype(np.int64)
    cols_t = torch.from_numpy(cols).cuda()
    nh = torch.empty(n_sample, dtype=torch.complex128, device="cuda")
    p64.ndft_cols_dev(p64.device_pointers["f"], cols_t.data_ptr(), n_sample, nh.data_ptr())
    torch.cuda.synchronize()
    nh = nh.cpu().numpy()
    out["adj_fp64_vs_ndft_rel_l2_sample"] = float(np.linalg.norm(h64[cols] - nh) / np.linalg.norm(nh))
    out["adj_fp32_vs_ndft_rel_l2_sample"] = float(np.linalg.norm(h32[cols] - nh) / np.linalg.norm(nh))
    return out


def measured_traffic(kern):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/), or None if it is absent."""
    path = os.path.join(ROOT, "profiles", "traffic_gridding_3d128_fp32.txt")
    key = {"fwd_gather_ms": "k_tile_gather", "adj_scatter_ms": "k_tile_spread"}[kern]
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) >= 2 and parts[0] == key + "_bytes_per_launch":
                    return float(parts[1])
    except OSError:
        pass
    return None


def cpu_baseline(x, fh, budget_s=15.0):
    """The oracle's FP64 direct NDFT (OpenMP, all host cores) on sampled rows of
    the same workload; throughput extrapolated linearly to the full fwd+adj NDFT."""
    import oracle

    oracle.build()
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = min(cores, 16)
    r = np.random.default_rng(11)
    n_rows = 8
    elapsed = 0.0
    while True:
        rows = r.choice(M_NODES, n_rows, replace=False).astype(np.int64)
        t0 = time.perf_counter()
        oracle.ndft_trafo(N3, x, fh, rows=rows, nthreads=cores)
        elapsed = time.perf_counter() - t0
        if elapsed > budget_s / 4 or n_rows >= 4096:
            break
        n_rows = int(n_rows * max(2.0, min(8.0, (budget_s / 4) / max(elapsed, 1e-3))))
    terms = n_rows * int(np.prod(N3))
    rate = terms / elapsed                              # terms per second
    full_terms = 2.0 * M_NODES * int(np.prod(N3))       # forward + adjoint NDFT
    value = M_NODES / (full_terms / rate)
    # one thread, on a smaller sample
    rows1 = rows[: max(1, n_rows // max(1, cores))]
    t0 = time.perf_counter()
    oracle.ndft_trafo(N3, x, fh, rows=rows1, nthreads=1)
    e1 = time.perf_counter() - t0
    value1 = M_NODES / (full_terms / (len(rows1) * int(np.prod(N3)) / e1))
    model = ""
    try:
        with open("/proc/cpuinfo") as fcpu:
            for line in fcpu:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": value, "unit": "nodes/s", "cores": int(cores), "kind": "port",
            "sample": f"FP64 direct NDFT (oracle, OpenMP) of {n_rows} sampled output rows of the "
                      f"forward transform ({terms:.3g} terms in {elapsed:.2f} s); fwd+adj rate "
                      f"extrapolated linearly to the full 2*M*N_total terms",
            "cores_note": "OpenMP threads used: the process's CPU affinity, capped at 16 "
                          "(the host share of one GPU)",
            "value_1_thread": value1, "sample_1_thread_rows": int(len(rows1)),
            "cpu_model": model, "nproc": os.cpu_count()}


BATCH_P, BATCH_R, BATCH_M = 96, 8, 4


def batch_workload(seed=5150):
    """The batched workload: 96 2D problems N_b x N_b, N_b drawn from
    {64, 96, 128, 192, 256}, sigma = 2 exactly (n_b = 2 N_b), m = 4, M_b
    uniform in [4096, 65536], 8 right-hand sides.  Returns (N, n, M,
    problem) with problem(i) -> (x (M_i, 2), fhat_in (8, N_i^2), f_in (8, M_i)),
    the coefficients complex64 with re/im N(0, 1); each problem has its own
    seeded stream, so any one can be rebuilt alone."""
    r = np.random.default_rng(seed)
    sizes = r.choice([64, 96, 128, 192, 256], BATCH_P)
    N = np.stack([sizes, sizes], axis=1)
    M = r.integers(4096, 65536, BATCH_P, endpoint=True)

    def problem(i):
        ri = np.random.default_rng([seed, i])
        Mi, Nt = int(M[i]), int(N[i][0] * N[i][1])
        x = ri.uniform(-0.5, 0.5, (Mi, 2))
        fh = (ri.standard_normal((BATCH_R, Nt)) + 1j * ri.standard_normal((BATCH_R, Nt))).astype(np.complex64)
        fv = (ri.standard_normal((BATCH_R, Mi)) + 1j * ri.standard_normal((BATCH_R, Mi))).astype(np.complex64)
        return x, fh, fv

    return N, 2 * N, M, problem


def load_batch(b, problem, dev):
    """Nodes and inputs of every problem into the batch b; returns the device
    node arrays (kept alive by the caller)."""
    import torch

    xs = []
    for i in range(b.P):
        x, fh, fv = problem(i)
        xt = torch.from_numpy(x).to(dev)
        xs.append(xt)
        b.set_nodes_dev(i, xt.data_ptr())
        fhat_in, _, f_in, _ = b.tensors(i)   # (R, N_total) and (R, M) views
        fhat_in.copy_(torch.from_numpy(fh).to(dev))
        f_in.copy_(torch.from_numpy(fv).to(dev))
    torch.cuda.synchronize()
    return xs


def bench_batch(cu, steps, warmup, seed=5150):
    """96 uneven 2D FP32 problems x 8 right-hand sides: one graph launch of the
    batched stage kernels per step, against a loop of single-plan launches."""
    import torch

    from cunfft_amd.batch import Batch

    P, R, m = BATCH_P, BATCH_R, BATCH_M
    N, n, M, problem = batch_workload(seed)
    sizes = N[:, 0]
    dev = torch.device("cuda")
    b = Batch(N, M, m=m, n_rhs=R, n=n, precision="single")   # sigma = 2 exactly
    xs = load_batch(b, problem, dev)
    config, ms_f, ms_a = b.autotune(reps=3)
    for _ in range(warmup):
        b.execute()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.execute()
    torch.cuda.synchronize()
    ms_batch = 1e3 * (time.perf_counter() - t0) / steps
    # the same problems as single plans, each autotuned, looped per right-hand side
    plans, ptrs = [], []
    for i in range(P):
        p = cu.Plan(list(N[i]), int(M[i]), n=list(b.n[i]), m=m, precision="single",
                    flags=cu.CUFFT_INIT | cu.CUFFT_BOTH | cu.NO_HOST_BUFFERS)
        p.set_nodes_dev(xs[i].data_ptr())
        p.autotune(reps=2)
        plans.append(p)
        fi, fo, ai, ao = b.buffers(i)
        Nt, Mi = int(N[i][0] * N[i][1]), int(M[i])
        ptrs.append([(fi + 8 * k * Nt, fo + 8 * k * Mi, ai + 8 * k * Mi, ao + 8 * k * Nt)
                     for k in range(R)])
    torch.cuda.synchronize()

    def loop_step():
        for p, pr in zip(plans, ptrs):
            for fi, fo, ai, ao in pr:
                p.transform_dev(fi, fo)
                p.adjoint_dev(ai, ao)

    for _ in range(warmup):
        loop_step()
    torch.cuda.synchronize()
    loop_steps = max(2, steps // 2)
    t0 = time.perf_counter()
    for _ in range(loop_steps):
        loop_step()
    torch.cuda.synchronize()
    ms_loop = 1e3 * (time.perf_counter() - t0) / loop_steps
    # accuracy: the batch's FP32 outputs against FP64 single plans on 3 problems
    b.execute()
    torch.cuda.synchronize()
    errs = []
    for i in (0, P // 2, P - 1):
        fi, fo, ai, ao = b.tensors(i)
        with cu.Plan(list(N[i]), int(M[i]), n=list(b.n[i]), m=m, precision="double",
                     flags=cu.CUFFT_INIT | cu.CUFFT_BOTH | cu.NO_HOST_BUFFERS) as p64:
            p64.set_nodes_dev(xs[i].data_ptr())
            f64 = torch.empty(int(M[i]), dtype=torch.complex128, device=dev)
            h64 = torch.empty(int(N[i][0] * N[i][1]), dtype=torch.complex128, device=dev)
            src_f = fi[0].to(torch.complex128).contiguous()
            src_a = ai[0].to(torch.complex128).contiguous()
            p64.transform_dev(src_f.data_ptr(), f64.data_ptr())
            p64.adjoint_dev(src_a.data_ptr(), h64.data_ptr())
            torch.cuda.synchronize()
            ef = float((fo[0].to(torch.complex128) - f64).norm() / f64.norm())
            eh = float((ao[0].to(torch.complex128) - h64).norm() / h64.norm())
            errs.append(max(ef, eh))
    for p in plans:
        p.finalize()
    b.close()
    nodes = int(M.sum()) * R
    return {
        "workload": "96 uneven 2D problems N_b^2, N_b in {64,96,128,192,256}, sigma=2 "
                    "(n_b = 2 N_b), m=4, M_b uniform in [4096,65536], 8 right-hand sides, "
                    "forward+adjoint, FP32",
        "value": nodes / (ms_batch * 1e-3), "unit": "nodes/s", "ms_per_step": ms_batch,
        "node_transforms_per_step": nodes, "loop_of_single_plans_ms_per_step": ms_loop,
        "loop_value": nodes / (ms_loop * 1e-3), "speedup_vs_loop": ms_loop / ms_batch,
        "batch_config": list(config), "autotune_fwd_ms": ms_f, "autotune_adj_ms": ms_a,
        "fp32_vs_fp64_plan_rel_l2_max_3_problems": max(errs),
        "n_b_counts": {int(v): int((sizes == v).sum()) for v in (64, 96, 128, 192, 256)},
    }


def config_sweep(cu, steps):
    """The FP32 headline workload (same nodes and data as the bench step) in each
    gridding configuration: `steps` forward and adjoint transforms per
    configuration after one warm-up, with the per-kernel event times of the last
    one.  Meant to run under a profiler, whose per-kernel counters then separate
    the configurations by kernel name (the tile shape is a template argument)."""
    import torch

    x, fh, fv = make_inputs()
    p = cu.Plan(N3, M_NODES, n=NN3, m=CUT, precision="single")
    p.x[:] = x
    p.f_hat[:] = fh
    p.f[:] = fv
    p.copyDataToDevice()
    p.copyDataToDeviceAd()
    dp = p.device_pointers
    out = {}
    for c in range(cu.TUNE_CONFIGS):
        try:
            p.config = c
        except cu.CunfftError as e:
            out[c] = f"refused: {e}"
            continue
        for _ in range(steps + 1):
            p.transform_dev(dp["f_hat"], dp["f"])
            p.adjoint_dev(dp["f"], dp["f_hat"])
        torch.cuda.synchronize()
        p._p.flags = p._p.flags | cu.MEASURE_TIME
        p.transform()
        tf = dict(p.times)
        p.adjoint()
        ta = dict(p.times)
        p._p.flags = p._p.flags & ~cu.MEASURE_TIME
        out[c] = {"forward": tf, "adjoint": ta}
    p.finalize()
    return {"config_sweep": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 2D workload")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched 2D workload (prints its JSON object)")
    ap.add_argument("--config-sweep", action="store_true",
                    help="only run the FP32 headline workload in every gridding configuration "
                         "(for per-configuration profiler counters) and print per-config times")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (to scale HBM "
                         "counters) and exit")
    a = ap.parse_args()
    if a.copy_calibration:
        import torch

        src = torch.empty(a.copy_calibration // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 4}))
        return 0
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the multi-GPU sharding is a design, not a run", file=sys.stderr)
        return 2
    import torch

    import cunfft_amd as cu

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if a.batch_only:
        print(json.dumps(bench_batch(cu, a.steps, a.warmup)), flush=True)
        return 0
    if a.config_sweep:
        print(json.dumps(config_sweep(cu, a.steps)), flush=True)
        return 0
    log("bench: building inputs")
    x, fh, fv = make_inputs()
    log("bench: timing FP32")
    p32, t32 = time_plan(cu, "single", x, fh, fv, a.steps, a.warmup)
    log(f"bench: FP32 {t32}")
    log("bench: timing FP64")
    p64, t64 = time_plan(cu, "double", x, fh, fv, a.steps, a.warmup)
    log(f"bench: FP64 {t64}")
    log("bench: accuracy")
    acc = accuracy(cu, p32, p64, x, fh)
    log(f"bench: {acc}")
    p32.finalize()
    p64.finalize()
    log("bench: PCIe-inclusive steps")
    pcie = pcie_inclusive(cu, x, fh, fv)
    log(f"bench: {pcie}")
    batch = None
    if not a.no_batch:
        log("bench: batched 2D workload")
        batch = bench_batch(cu, a.steps, a.warmup)
        log(f"bench: {batch}")

    value = M_NODES / (t32["ms_per_step"] * 1e-3)
    value64 = M_NODES / (t64["ms_per_step"] * 1e-3)
    # dominant kernel of the FP32 step and its roofline (algorithmic bytes of one
    # gridding launch: the grid once plus the nodes, see grid_bytes)
    cf, ca = t32["config"]
    shapes = ["8x8x8", "8x8x16", "16x16x8"]
    names = {"fwd_gather_ms": f"k_tile_gather<float,3,10,NW,{cf}> ({shapes[cf]} tiles, forward "
                              "gridding)" if cf < 3 else "k_gather_tiled<float,3,10> (node lanes)",
             "adj_scatter_ms": "k_tile_spread_mfma<float,10,0> (8x8x8 tiles, adjoint tile accumulation on "
                               "the matrix cores)" if ca == 0
                               else f"k_tile_spread_queue<float,3,10,4,{ca}> ({shapes[ca]} tiles, "
                               "adjoint tile accumulation)"
                               if ca < 3 else "k_scatter_tiled<float,double,3,10> (node lanes)"}
    kern = max(names, key=lambda k: t32[k])
    ach = grid_bytes(8) / (t32[kern] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": names[kern], "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": measured_traffic(kern) if (cf if kern == "fwd_gather_ms" else ca) == 0
            else None,
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of this kernel in this "
                              "configuration, scaled by a plain device copy's counters "
                              "(profiles/traffic_gridding_3d128_fp32.txt); null when that "
                              "summary does not cover the configuration the autotuner chose",
            "algorithmic_bytes": grid_bytes(8), "avg_launch_ms": t32[kern]}
    # the spread's products on the matrix cores: the useful flops (two FMAs per window
    # cell and node, complex coefficient) against the dense MFMA peak of the precision
    useful = M_NODES * (2 * CUT + 2) ** 3 * 4

    def matrix_cores(ms, peak):
        return {"useful_flops": useful, "useful_tflops": useful / (ms * 1e-3) / 1e12,
                "peak_tflops": peak, "frac": useful / (ms * 1e-3) / 1e12 / peak}

    if kern == "adj_scatter_ms" and ca == 0:
        roof["matrix_cores"] = matrix_cores(t32[kern], MFMA_FP32_PEAK_TFLOPS)
    fp64 = {"value": value64, "ms_per_step": t64["ms_per_step"], "fp32_speedup_over_fp64": value / value64}
    if t64["config"][1] == 0:
        fp64["adj_kernel"] = ("k_tile_spread_mfma<double,10,0> (8x8x8 tiles, adjoint tile accumulation "
                              "on the FP64 matrix cores)")
        fp64["adj_matrix_cores"] = matrix_cores(t64["adj_scatter_ms"], MFMA_FP64_PEAK_TFLOPS)
    cpu = None
    if not a.no_cpu_baseline:
        log("bench: CPU baseline")
        cpu = cpu_baseline(x, fh)
    line = {
        "metric": "fwd+adj NFFT nodes/s, 3D 128^3 M=2^21, FP32 vs FP64 @<=1e-5 rel err; %HBM peak",
        "value": value, "unit": "nodes/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": t32["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64 (fp32 complex; fp64 index split and roll-off)",
        "data": "synthetic: half uniform, half N(0,0.1^2) wrapped nodes; N(0,1) coefficients",
        "config": {"workload": "3D NFFT forward+adjoint, N=128^3, n=256^3, m=4, M=2^21 nodes",
                   "N": N3, "n": NN3, "m": CUT, "M": M_NODES, "precision": "fp32"},
        "fp64": fp64,
        "steps_fp32_ms": t32, "steps_fp64_ms": t64, "accuracy": acc,
        "roofline": roof, "cpu_baseline": cpu, "batched_2d": batch, "pcie_inclusive": pcie,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
