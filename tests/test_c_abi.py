"""The C boundary without a GPU: the header compiles as C for both precision
families, the plan layout a C compiler sees matches the ctypes mirror, and
the host-only entry points (setProperties, nfft_init_fg_exp_l[_m],
resetCUNFFTTimes, the harness of cunfft_util.h) compute what the reference's
do (src/cunfft.cu:31-90, src/cuda_util.cpp:151-258).
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
CDIR = os.path.join(ROOT, "tests", "c_abi")


def lib():
    import cunfft_amd

    if not os.path.exists(cunfft_amd.LIB_PATH):
        from cunfft_amd import build

        build.build(force=False)
    return ctypes.CDLL(cunfft_amd.LIB_PATH)


@pytest.mark.parametrize("single", [False, True])
def test_tutorial_compiles_as_c(tmp_path, single):
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-D_POSIX_C_SOURCE=199309L",
           "-I", INC, "-c", os.path.join(CDIR, "tutorial.c"), "-o", str(tmp_path / "t.o")]
    if single:
        cmd.insert(1, "-DCUNFFT_SINGLE_PRECISION")
    subprocess.run(cmd, check=True)


def test_plan_layout_matches_the_ctypes_mirror(tmp_path):
    from cunfft_amd import cunfft_plan, NFFTTimeSpec

    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INC,
                    os.path.join(CDIR, "layout.c"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in out:
        parts = line.split()
        if len(parts) != 3:
            continue
        typ, field, val = parts[0], parts[1], int(parts[2])
        if typ in ("cunfft_plan", "cunfftf_plan"):
            if field == "sizeof":
                assert ctypes.sizeof(cunfft_plan) == val, (typ, field)
            else:
                assert getattr(cunfft_plan, field).offset == val, (typ, field)
            seen += 1
        elif typ == "NFFTTimeSpec":
            assert ctypes.sizeof(NFFTTimeSpec) == val
    assert seen == 2 * 26


def test_set_properties_and_fg_table():
    L = lib()
    d = 3
    N = (ctypes.c_ssize_t * 3)(64, 32, 16)
    n = (ctypes.c_ssize_t * 3)(128, 64, 48)
    sigma = (ctypes.c_double * 3)(*[n[t] / N[t] for t in range(d)])
    m = 6
    b = (ctypes.c_double * 3)(*[(2 * s) / (2 * s - 1) * m / math.pi for s in sigma])

    class KData(ctypes.Structure):
        _fields_ = [(k, v * 3) for k, v in (("N", ctypes.c_ssize_t), ("n", ctypes.c_ssize_t),
                                             ("b", ctypes.c_double), ("bsqrt", ctypes.c_double),
                                             ("fac", ctypes.c_double), ("sigma", ctypes.c_double))]

    k = KData()
    L.setProperties(N, n, b, sigma, d, ctypes.byref(k))
    for t in range(d):
        assert k.N[t] == N[t] and k.n[t] == n[t] and k.b[t] == b[t] and k.sigma[t] == sigma[t]
        assert abs(k.bsqrt[t] - math.sqrt(b[t])) <= 1e-15
        assert abs(k.fac[t] - math.pi ** 2 * b[t] / n[t] ** 2) <= 1e-15 * k.fac[t]
    # src/cunfft.cu:31-76: exp(-l^2 / b) for l = 0 .. 2m+1 at the offset
    tab = (ctypes.c_double * 40)()
    L.nfft_init_fg_exp_l.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    L.nfft_init_fg_exp_l(tab, b[0], 3)
    for l in range(2 * 6 + 2):
        assert abs(tab[3 + l] - math.exp(-l * l / b[0])) <= 1e-15
    assert tab[2] == 0.0 and tab[3 + 14] == 0.0
    L.nfft_init_fg_exp_l_m.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    tab2 = (ctypes.c_double * 40)()
    L.nfft_init_fg_exp_l_m(tab2, 2.5, 0, 4)
    assert [tab2[l] for l in range(10)] == pytest.approx([math.exp(-l * l / 2.5) for l in range(10)], rel=1e-15)
    assert tab2[10] == 0.0


def test_reset_times_and_math_helpers():
    from cunfft_amd import NFFTTimeSpec

    L = lib()
    t = NFFTTimeSpec(*([1.5] * 7))
    L.resetCUNFFTTimes(ctypes.byref(t))
    assert all(getattr(t, f) == 0.0 for f, _ in NFFTTimeSpec._fields_)
    L.prod_int.restype = ctypes.c_ssize_t
    L.next_power_of_2.restype = ctypes.c_ssize_t
    L.next_power_of_2.argtypes = [ctypes.c_ssize_t]
    v = (ctypes.c_ssize_t * 3)(3, 5, 7)
    assert L.prod_int(v, 3) == 105
    assert [L.next_power_of_2(x) for x in (1, 2, 3, 1000, 1024)] == [1, 2, 4, 1024, 1024]


def test_error_norms_match_the_reference_definition():
    L = lib()
    r = np.random.default_rng(1)
    a = r.standard_normal(50) + 1j * r.standard_normal(50)
    b = a + 1e-3 * (r.standard_normal(50) + 1j * r.standard_normal(50))
    for fn in (L.l_infty, L.compute_error_l_infty):
        fn.restype = ctypes.c_double
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    pa, pb = a.ctypes.data, b.ctypes.data
    assert L.l_infty(pa, pb, 50) == pytest.approx(np.abs(a - b).max(), rel=1e-15)
    assert L.l_infty(pa, None, 50) == pytest.approx(np.abs(a).max(), rel=1e-15)
    assert L.compute_error_l_infty(pa, pb, 50) == pytest.approx(np.abs(a - b).max() / np.abs(a).max(), rel=1e-15)


def test_example_data_is_the_drand48_stream(orc):
    """getExampleData_uniDistr / getExampleDataAd_uniDistr draw from the C
    library's drand48: from state 0 (seed48 of zeros, an unseeded process)
    they give the oracle's restatement bit for bit."""
    L = lib()
    libc = ctypes.CDLL(None)
    libc.seed48.argtypes = [ctypes.POINTER(ctypes.c_ushort)]
    libc.seed48.restype = ctypes.c_void_p
    L.createFourierCoeff_uniDistr.argtypes = [ctypes.c_void_p, ctypes.c_ssize_t]
    zero = (ctypes.c_ushort * 3)(0, 0, 0)
    libc.seed48(zero)
    fh = np.zeros(37, np.complex128)
    L.createFourierCoeff_uniDistr(fh.ctypes.data, 37)
    r = orc.Drand48(0)
    ref = np.array([complex(r.next(), r.next()) for _ in range(37)])
    assert np.array_equal(fh, ref)
