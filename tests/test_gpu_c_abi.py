"""A C program built against include/cunfft.h + cunfft_util.h and linked
with libcunfft_amd.so runs the reference's tutorial sequence
(src/simpleTest.cpp:18-148): cunfft_init, getExampleData_uniDistr, CUNDFT,
CUNFFT after cunfft_reinit, CUNDFT and CUNFFT adjoints after cunfft_reinitAd
(taking the forward result as input), for a 1D and a 2D plan, in FP64 and
with CUNFFT_SINGLE_PRECISION.  Its inputs must be the drand48 stream from
state 0 and its outputs the oracle's on those inputs.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
CDIR = os.path.join(ROOT, "tests", "c_abi")
LIBDIR = os.path.join(ROOT, "cunfft_amd")


def build_and_run(tmp_path, single):
    exe = tmp_path / ("tutorial_f32" if single else "tutorial_f64")
    cmd = ["gcc", "-std=c99", "-O1", "-D_POSIX_C_SOURCE=199309L", "-I", INC,
           os.path.join(CDIR, "tutorial.c"), "-o", str(exe), "-L", LIBDIR, "-lcunfft_amd",
           "-Wl,-rpath," + LIBDIR, "-lm"]
    if single:
        cmd.insert(1, "-DCUNFFT_SINGLE_PRECISION")
    subprocess.run(cmd, check=True)
    out = tmp_path / ("out32" if single else "out64")
    out.mkdir()
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "TUTORIAL_OK" in r.stdout
    return out


def load(out, d, name, single, kind="c"):
    a = np.fromfile(out / f"{d}d_{name}.bin", dtype=np.float32 if single and kind == "c" else np.float64)
    return a.view(np.complex64 if single else np.complex128).astype(np.complex128) if kind == "c" else a


@pytest.mark.parametrize("single", [False, True])
def test_tutorial_in_c_matches_the_oracle(gpu, orc, tmp_path, single):
    out = build_and_run(tmp_path, single)
    r = orc.Drand48(0)
    tol = 1e-5 if single else 1e-12
    for d, N in ((1, [32]), (2, [32, 32])):
        M = int(np.prod(N))
        n = orc.default_n(N).tolist()
        m = 6   # cunfft_init's cut-off (CMakeLists.txt:170)
        fh_ref, x_ref = r.example_data(N, M)
        x = load(out, d, "x", single, "r").reshape(M, d)
        fh = load(out, d, "fhat_in", single)
        assert np.array_equal(x, x_ref)
        q = fh_ref.astype(np.complex64).astype(np.complex128) if single else fh_ref
        assert np.array_equal(fh, q)
        f_ndft = load(out, d, "f_ndft", single)
        f_nfft = load(out, d, "f_nfft", single)
        e1 = orc.rel_l2(orc.ndft_trafo(N, x, fh), f_ndft)
        e2 = orc.rel_l2(orc.nfft_trafo(N, n, m, x, fh), f_nfft)
        h_ndft = load(out, d, "fhat_ndft", single)
        h_nfft = load(out, d, "fhat_nfft", single)
        e3 = orc.rel_l2(orc.ndft_adjoint(N, x, f_nfft), h_ndft)
        e4 = orc.rel_l2(orc.nfft_adjoint(N, n, m, x, f_nfft), h_nfft)
        print(d, "single" if single else "double", "NDFT", e1, "NFFT", e2, "NDFT adj", e3, "NFFT adj", e4)
        assert max(e1, e2, e3, e4) <= tol
        # the window error of m = 6 between the two transforms
        assert orc.rel_l2(f_ndft, f_nfft) <= 1e-6


def test_warmup_and_copy_x_to_host(gpu):
    L = gpu.lib()
    n = (ctypes.c_ssize_t * 3)(64, 64, 64)
    L.warmUp_createPlan.argtypes = [ctypes.c_void_p]
    L.warmUp_createPlan(n)
    r = np.random.default_rng(3)
    x = r.uniform(-0.5, 0.5, (300, 2))
    with gpu.Plan([16, 16], 300) as p:
        p.x[:] = x
        p.copyDataToDevice()
        assert np.array_equal(p.copy_x_toHost(), x)
