"""GPU parity of the HIP path against the CPU oracle and the golden fixtures.

Every call goes through the C ABI of libcunfft_amd.so (cunfft_amd.Plan is a
ctypes mirror of the reference's plan surface).  Tolerances:
  FP64 plans vs the FP64 oracle NFFT   rel-l2 <= 1e-12 (same algorithm, roundoff only)
  FP32 plans vs the FP64 oracle NFFT   rel-l2 <= 1e-5  (precision error on the same window)
  any plan vs the NDFT                 the window band of the oracle tests
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "golden_cases.npz"))
GOLD_CASES = sorted({k.split("/")[0] for k in GOLD.files})

TOL = {"double": 1e-12, "single": 1e-5}


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def run_forward(cu, N, n, m, x, fh, precision, flags=None):
    with cu.Plan(N, x.shape[0], n=n, m=m, precision=precision, flags=flags) as p:
        p.x[:] = x
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        return p.f.astype(np.complex128).copy()


def run_adjoint(cu, N, n, m, x, fv, precision, flags=None):
    with cu.Plan(N, x.shape[0], n=n, m=m, precision=precision, flags=flags) as p:
        p.x[:] = x
        p.f[:] = fv
        p.copyDataToDeviceAd()
        p.adjoint()
        p.copyDataToHostAd()
        return p.f_hat.astype(np.complex128).copy()


def q(a, precision):
    """Inputs as the plan stores them (FP32 plans round f_hat / f once)."""
    return a.astype(np.complex64).astype(np.complex128) if precision == "single" else a


# ------------------------------------------------------------ golden cases --
@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("case", GOLD_CASES)
def test_golden_forward_adjoint(gpu, orc, case, precision):
    g = {k.split("/")[1]: GOLD[k] for k in GOLD.files if k.startswith(case + "/")}
    N, n, m = list(g["N"]), list(g["n"]), int(g["m"])
    f = run_forward(gpu, N, n, m, g["x"], g["fhat"], precision)
    h = run_adjoint(gpu, N, n, m, g["x"], g["f"], precision)
    if precision == "double":
        ref_f, ref_h = g["nfft_trafo"], g["nfft_adjoint"]
    else:
        ref_f = orc.nfft_trafo(N, n, m, g["x"], q(g["fhat"], precision))
        ref_h = orc.nfft_adjoint(N, n, m, g["x"], q(g["f"], precision))
    print(case, precision, orc.rel_l2(ref_f, f), orc.rel_l2(ref_h, h))
    assert orc.rel_l2(ref_f, f) <= TOL[precision]
    assert orc.rel_l2(ref_h, h) <= TOL[precision]
    # and the NDFT within the window band (m = 2 is the widest band)
    band = {2: 1.5e-2, 3: 2e-3, 4: 3.5e-4, 6: 5e-6}[m]
    assert orc.rel_l2(g["ndft_trafo"], f) <= band
    assert orc.rel_l2(g["ndft_adjoint"], h) <= band


# --------------------------------------------------- per-stage and sweeps --
SWEEP = [
    ([1024], None, 6), ([64], None, 1), ([256], None, 8), ([48], [128], 5),
    ([64, 64], None, 4), ([32, 16], [64, 64], 3), ([96, 40], [256, 128], 4),
    ([16, 16, 16], None, 4), ([8, 16, 32], None, 3), ([12, 8, 20], [32, 16, 64], 2),
    ([32], [64], 9), ([64], None, 16), ([16, 16], None, 9), ([8, 8, 8], [32, 32, 32], 10),
    ([32, 32], [32, 32], 4),   # sigma = 1
]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N,n,m", SWEEP)
def test_sweep_matches_oracle(gpu, orc, N, n, m, precision):
    r = np.random.default_rng(hash((tuple(N), m)) % 2 ** 32)
    n = n or orc.default_n(N).tolist()
    M = 517
    x = r.uniform(-0.5, 0.5, (M, len(N)))
    fh, fv = crandn(r, int(np.prod(N))), crandn(r, M)
    f = run_forward(gpu, N, n, m, x, fh, precision)
    h = run_adjoint(gpu, N, n, m, x, fv, precision)
    ef = orc.rel_l2(orc.nfft_trafo(N, n, m, x, q(fh, precision)), f)
    eh = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, q(fv, precision)), h)
    print(N, n, m, precision, ef, eh)
    # at sigma = 1 the roll-off factors reach exp(pi m / 2) per axis (5e5 at the 2D
    # corner for m = 4): the grid is that much larger than the result, so roundoff
    # grows by the same conditioning factor (FP64 1e-12 -> 2e-11, FP32 1e-5 -> 1e-3)
    tol = TOL[precision] * (20 if precision == "double" else 100) \
        if min(a / b for a, b in zip(n, N)) < 1.5 else TOL[precision]
    assert ef <= tol and eh <= tol


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N", [[2048], [128, 256], [32, 32, 32]])
def test_fft_stage_grid(gpu, orc, N, precision):
    """The grid after the forward FFT passes equals FFT(roll-off/pad(fhat))."""
    r = np.random.default_rng(3)
    n = orc.default_n(N).tolist()
    m = 4
    fh = crandn(r, int(np.prod(N)))
    with gpu.Plan(N, 10, n=n, m=m, precision=precision) as p:
        p.x[:] = r.uniform(-0.5, 0.5, (10, len(N)))
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copy_g_ToHost()
        g = p.g.astype(np.complex128)
    ref = orc.fft(n, orc.rolloff_pad(N, n, m, q(fh, precision)), -1)
    e = orc.rel_l2(ref, g)
    print(N, precision, e)
    assert e <= {"double": 1e-13, "single": 2e-6}[precision]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("N", [[100], [24, 16], [8, 6, 10]])
def test_gpu_ndft_matches_cpu_ndft(gpu, orc, N, precision):
    r = np.random.default_rng(4)
    M = 300
    x = r.uniform(-0.5, 0.5, (M, len(N)))
    fh, fv = q(crandn(r, int(np.prod(N))), precision), q(crandn(r, M), precision)
    with gpu.Plan(N, M, precision=precision) as p:
        p.x[:] = x
        p.f_hat[:] = fh
        p.f[:] = fv
        p.copyDataToDevice()
        p.cundft_transform()          # writes f_gpu, as src/cunfft.cu:426-431
        f = p.copy_f_ToHost()
        p.copyDataToDeviceAd()
        p.cundft_adjoint()            # writes fhat_gpu
        h = p.copy_f_hat_ToHost()
    # FP32 plans sum in FP64 and round the result once
    tol = 1e-13 if precision == "double" else 1e-7
    ef = orc.rel_l2(orc.ndft_trafo(N, x, fh), f)
    eh = orc.rel_l2(orc.ndft_adjoint(N, x, fv), h)
    assert ef <= tol and eh <= tol, (ef, eh)
    assert np.array_equal(p_x := x, x) and f.dtype == (np.complex128 if precision == "double"
                                                         else np.complex64)


# ------------------------------------------------------------- edge cases --
@pytest.mark.parametrize("precision", ["double", "single"])
def test_edge_nodes(gpu, orc, precision):
    N, m = [32, 32], 4
    n = orc.default_n(N).tolist()
    x = np.array([[-0.5, -0.5], [0.0, 0.0], [-0.0, 0.0], [0.0, -0.0], [-0.0, -0.0],
                  [0.49999999999, -0.5], [0.5 - 2 ** -40, 0.25], [1 / 64, -1 / 64],
                  [0.5 / 64, -0.5 / 64], [1.5 / 64, -2.5 / 64]])
    fh = crandn(np.random.default_rng(5), 1024)
    f = run_forward(gpu, N, n, m, x, fh, precision)
    assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, q(fh, precision)), f) <= TOL[precision]


@pytest.mark.parametrize("precision", ["double", "single"])
def test_single_node_and_coincident_nodes(gpu, orc, precision):
    N, m = [16, 16, 16], 4
    n = orc.default_n(N).tolist()
    r = np.random.default_rng(6)
    for M, x in [(1, np.array([[0.1, -0.2, 0.3]])),
                 (700, np.tile([[0.0, 0.0, 0.0]], (700, 1))),
                 (700, np.clip(r.normal(0, 0.01, (700, 3)), -0.5, 0.49))]:
        fv = crandn(r, M)
        h = run_adjoint(gpu, N, n, m, x, fv, precision)
        ref = orc.nfft_adjoint(N, n, m, x, q(fv, precision))
        assert orc.rel_l2(ref, h) <= TOL[precision] * 10


@pytest.mark.parametrize("N", [[32, 32, 32], [64, 128]])
@pytest.mark.parametrize("flags_name", ["ATOMIC_SCATTER", "PLAIN_GATHER", "NODE_LANES"])
def test_plain_paths(gpu, orc, flags_name, N):
    m = 4
    n = orc.default_n(N).tolist()
    r = np.random.default_rng(7)
    x = r.uniform(-0.5, 0.5, (3000, len(N)))
    fh, fv = crandn(r, int(np.prod(N))), crandn(r, 3000)
    flags = gpu.CUFFT_INIT | gpu.CUFFT_BOTH | getattr(gpu, flags_name)
    f = run_forward(gpu, N, n, m, x, fh, "double", flags)
    h = run_adjoint(gpu, N, n, m, x, fv, "double", flags)
    assert orc.rel_l2(orc.nfft_trafo(N, n, m, x, fh), f) <= 1e-12
    assert orc.rel_l2(orc.nfft_adjoint(N, n, m, x, fv), h) <= 1e-12


# ---------------------------------------------------------- plan lifecycle --
def test_reinit_and_repeat(gpu, orc):
    N, m = [64, 32], 4
    n = orc.default_n(N).tolist()
    r = np.random.default_rng(8)
    M = 1000
    x = r.uniform(-0.5, 0.5, (M, 2))
    fh = crandn(r, 2048)
    with gpu.Plan(N, M, n=n, m=m) as p:
        p.x[:] = x
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        f1 = p.f.copy()
        p.reinit()
        assert not np.any(p.f)
        p.copyDataToDevice()
        p.transform()
        p.transform()
        p.copyDataToHost()
        assert np.array_equal(f1, p.f)
        p.reinitAd()
        assert not np.any(p.f_hat)
        # new nodes: binning must be rebuilt
        x2 = r.uniform(-0.5, 0.5, (M, 2))
        p.x[:] = x2
        p.f_hat[:] = fh
        p.copyDataToDevice()
        p.transform()
        p.copyDataToHost()
        assert orc.rel_l2(orc.nfft_trafo(N, n, m, x2, fh), p.f) <= 1e-12


def test_measured_times(gpu):
    with gpu.Plan([64, 64], 5000, flags=gpu.CUFFT_INIT | gpu.CUFFT_BOTH | gpu.MEASURE_TIME) as p:
        p.x[:] = np.random.default_rng(9).uniform(-0.5, 0.5, (5000, 2))
        p.copyDataToDevice()
        p.transform()
        t = p.times
        assert t["time_FFT"] > 0 and t["time_CONV"] > 0 and t["time_COPY_IN"] > 0
        assert t["runTime"] == pytest.approx(t["time_FFT"] + t["time_CONV"])


def test_errors_do_not_crash(gpu):
    cu = gpu
    for kw in [dict(N=[30], M=10, n=[40]), dict(N=[32], M=0), dict(N=[32], M=10, m=0),
               dict(N=[8], M=10, n=[8], m=4), dict(N=[32], M=10, n=[2 ** 25]),
               dict(N=[32], M=10, n=[8200])]:
        with pytest.raises(cu.CunfftError):
            cu.Plan(kw.pop("N"), kw.pop("M"), **kw)
    p = cu.Plan([16], 10)
    p.finalize()
    p.finalize()  # second finalize is a no-op
    with pytest.raises(cu.CunfftError) as e:
        p.transform()
    assert e.value.status == 5
    p2 = cu.Plan([16], 10, flags=cu.CUFFT_INIT | cu.CUFFT_BOTH | cu.NO_HOST_BUFFERS)
    with pytest.raises(cu.CunfftError):
        p2.copyDataToDevice()
    p2.finalize()
