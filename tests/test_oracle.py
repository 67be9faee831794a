"""Pins the CPU oracle (oracle/liboracle.so) before anything is checked against it.

Every check here is independent of the oracle's own formulas: glibc's drand48,
numpy's FFT, closed forms of the NDFT, long-double brute force, the
adjointness identity, finite differences, and the error bands the reference
recorded for its own FP64 runs (Tests/Runtime/*.txt).  The planted-mistake
tests show that a plausible error in the pipeline fails these checks.
"""
import ctypes
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def rng(seed):
    return np.random.default_rng(seed)


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


# ------------------------------------------------------------------ inputs --
def test_drand48_matches_glibc(orc):
    libc = ctypes.CDLL("libc.so.6")
    libc.erand48.restype = ctypes.c_double
    libc.erand48.argtypes = [ctypes.POINTER(ctypes.c_ushort)]
    xs = (ctypes.c_ushort * 3)(0, 0, 0)  # unseeded process state
    ours = orc.Drand48(0)
    for _ in range(2000):
        assert ours.next() == libc.erand48(xs)


def test_example_data_order(orc):
    # getExampleData_uniDistr: fhat re, im for every k, then x - 0.5 (src/cuda_util.cpp:151-212)
    r = orc.Drand48(0)
    fh, x = r.example_data([4, 2], 3)
    s = orc.Drand48(0)
    seq = [s.next() for _ in range(16 + 6)]
    assert np.array_equal(fh.real, seq[0:16:2]) and np.array_equal(fh.imag, seq[1:16:2])
    assert np.array_equal(x.reshape(-1), np.array(seq[16:]) - 0.5)


# ------------------------------------------------------- window and indices --
def test_shape_parameter_and_window(orc):
    m = 6
    b = orc.shape_b(64, 128, m)
    assert b == pytest.approx(4.0 / 3.0 * m / np.pi, rel=1e-15)
    # phi integrates to 1 and peaks at (pi b)^-1/2
    t = np.linspace(-40, 40, 160001)
    vals = np.array([orc.phi(v, b) for v in t[::10]])
    assert np.trapezoid(vals, t[::10]) == pytest.approx(1.0, rel=1e-10)
    assert orc.phi(0.0, b) == pytest.approx(1 / np.sqrt(np.pi * b), rel=1e-15)


@pytest.mark.parametrize("n,x,m,u", [
    (16, 0.0, 4, -4),            # +0.0: window [-m, m+1]
    (16, -0.0, 4, -5),           # -0.0: signbit shifts the window to [-m-1, m]
    (16, 2.5 / 16, 4, 2 - 4),    # n x = 2.5 rounds half to even -> 2
    (16, 3.5 / 16, 4, 4 - 4),    # n x = 3.5 -> 4
    (16, -2.5 / 16, 4, -2 - 1 - 4),
    (16, -0.5, 3, -8 - 1 - 3),
    (256, 0.49999, 4, 128 - 4),
])
def test_window_start_rule(orc, n, x, m, u):
    assert orc.window_start(n, x, m) == u


def test_window_covers_neighbourhood(orc):
    # every grid point within distance m of n x is inside [u, u+2m+1]
    r = rng(3)
    for _ in range(2000):
        n, m = 64, int(r.integers(1, 9))
        x = r.uniform(-0.5, 0.5)
        u = orc.window_start(n, x, m)
        lo, hi = n * x - m, n * x + m
        assert u <= np.ceil(lo) and np.floor(hi) <= u + 2 * m + 1


# --------------------------------------------------------------------- FFT --
@pytest.mark.parametrize("n", [[8], [64], [2048], [16, 32], [8, 4, 16],
                               [6], [48], [192], [24, 96], [12, 6, 18]])
def test_fft_matches_numpy(orc, n):
    g = crandn(rng(1), int(np.prod(n)))
    fw = orc.fft(n, g, -1)
    bw = orc.fft(n, g, +1)
    ref_fw = np.fft.fftn(g.reshape(n)).reshape(-1)
    ref_bw = np.fft.ifftn(g.reshape(n)).reshape(-1) * np.prod(n)
    assert np.max(np.abs(fw - ref_fw)) <= 1e-12 * np.max(np.abs(ref_fw))
    assert np.max(np.abs(bw - ref_bw)) <= 1e-12 * np.max(np.abs(ref_bw))


def test_rolloff_positions_and_factors(orc):
    N, n, m = [6, 4], [16, 8], 3
    b = [orc.shape_b(N[t], n[t], m) for t in range(2)]
    fac = [np.pi ** 2 * b[t] / n[t] ** 2 for t in range(2)]
    for kl in range(24):
        fh = np.zeros(24, complex)
        fh[kl] = 1 + 2j
        k0, k1 = kl // 4 - 3, kl % 4 - 2
        g = orc.rolloff_pad(N, n, m, fh).reshape(n)
        expect = (1 + 2j) * np.exp(fac[0] * k0 ** 2 + fac[1] * k1 ** 2)
        assert g[k0 % 16, k1 % 8] == pytest.approx(expect, rel=1e-14)
        assert np.count_nonzero(g) == 1
        back = orc.crop_rolloff(N, n, m, g.reshape(-1))
        assert back[kl] == pytest.approx(expect * np.exp(fac[0] * k0 ** 2 + fac[1] * k1 ** 2), rel=1e-14)
        assert np.count_nonzero(back) == 1


# -------------------------------------------------------------------- NDFT --
def test_ndft_unit_coefficient(orc):
    N = [8, 6]
    x = rng(2).uniform(-0.5, 0.5, (50, 2))
    for kl in (0, 7, 25, 47):
        fh = np.zeros(48, complex)
        fh[kl] = 1.0
        k = np.array([kl // 6 - 4, kl % 6 - 3])
        f = orc.ndft_trafo(N, x, fh)
        assert np.allclose(f, np.exp(-2j * np.pi * x @ k), atol=1e-14)


def test_ndft_origin_and_equispaced(orc):
    N = [16]
    fh = crandn(rng(4), 16)
    assert np.allclose(orc.ndft_trafo(N, np.zeros((3, 1)), fh), fh.sum(), atol=1e-13)
    # x_j = -1/2 + j/N reduces to a DFT of (-1)^k fhat (in FFT order)
    x = (-0.5 + np.arange(16) / 16).reshape(-1, 1)
    k = np.arange(16) - 8
    dft = np.array([np.sum(fh * (-1.0) ** k * np.exp(-2j * np.pi * k * j / 16)) for j in range(16)])
    assert np.allclose(orc.ndft_trafo(N, x, fh), dft, atol=1e-12)
    via_fft = np.fft.fft(np.fft.ifftshift(fh * (-1.0) ** k))
    assert np.allclose(dft, via_fft, atol=1e-12)


def test_ndft_long_double_brute_force(orc):
    N = [4, 4, 2]
    r = rng(5)
    x = r.uniform(-0.5, 0.5, (7, 3))
    fh = crandn(r, 32)
    ks = np.array([[a - 2, b - 2, c - 1] for a in range(4) for b in range(4) for c in range(2)],
                  dtype=np.longdouble)
    xl = x.astype(np.longdouble)
    ph = -2 * np.pi * (xl @ ks.T)
    fr = fh.real.astype(np.longdouble)
    fi = fh.imag.astype(np.longdouble)
    re = (np.cos(ph) * fr - np.sin(ph) * fi).sum(axis=1)
    im = (np.sin(ph) * fr + np.cos(ph) * fi).sum(axis=1)
    f = orc.ndft_trafo(N, x, fh)
    assert np.max(np.abs(f.real - re.astype(float))) < 1e-13
    assert np.max(np.abs(f.imag - im.astype(float))) < 1e-13
    # adjoint brute force
    fv = crandn(r, 7)
    fr, fi = fv.real.astype(np.longdouble), fv.imag.astype(np.longdouble)
    re = (np.cos(-ph) * fr[:, None] - np.sin(-ph) * fi[:, None]).sum(axis=0)
    im = (np.sin(-ph) * fr[:, None] + np.cos(-ph) * fi[:, None]).sum(axis=0)
    h = orc.ndft_adjoint(N, x, fv)
    assert np.max(np.abs(h.real - re.astype(float))) < 1e-13
    assert np.max(np.abs(h.imag - im.astype(float))) < 1e-13


def test_ndft_rows_and_cols_subsets(orc):
    N = [8, 8]
    r = rng(6)
    x = r.uniform(-0.5, 0.5, (40, 2))
    fh, fv = crandn(r, 64), crandn(r, 40)
    rows = np.array([3, 0, 39, 17])
    assert np.allclose(orc.ndft_trafo(N, x, fh, rows=rows), orc.ndft_trafo(N, x, fh)[rows], rtol=0, atol=0)
    cols = np.array([63, 5, 8])
    assert np.allclose(orc.ndft_adjoint(N, x, fv, cols=cols), orc.ndft_adjoint(N, x, fv)[cols], rtol=0, atol=0)


# ------------------------------------------------------------------- NFFT --
CASES = [  # (N, m, band on rel-l2 vs NDFT) -- the Gaussian window's own error
    ([64], 2, 1.5e-2), ([64], 4, 2e-4), ([128], 6, 3e-6), ([64], 8, 3e-8), ([64], 12, 6e-12),
    ([32, 32], 4, 3e-4), ([16, 24], 5, 2.5e-5), ([32, 32], 6, 3e-6),
    ([16, 16, 16], 4, 3.5e-4), ([8, 12, 16], 6, 5e-6),
]


@pytest.mark.parametrize("N,m,band", CASES)
def test_nfft_matches_ndft_within_window_band(orc, N, m, band):
    r = rng(7)
    n = orc.default_n(N)
    M = 400
    x = r.uniform(-0.5, 0.5, (M, len(N)))
    fh, fv = crandn(r, int(np.prod(N))), crandn(r, M)
    e_f = orc.rel_l2(orc.ndft_trafo(N, x, fh), orc.nfft_trafo(N, n, m, x, fh))
    e_a = orc.rel_l2(orc.ndft_adjoint(N, x, fv), orc.nfft_adjoint(N, n, m, x, fv))
    assert e_f < band and e_a < band


@pytest.mark.parametrize("N,m", [([64], 4), ([16, 8], 3), ([8, 8, 8], 4)])
def test_nfft_adjointness(orc, N, m):
    r = rng(8)
    n = orc.default_n(N)
    x = r.uniform(-0.5, 0.5, (300, len(N)))
    fh, fv = crandn(r, int(np.prod(N))), crandn(r, 300)
    lhs = np.vdot(fv, orc.nfft_trafo(N, n, m, x, fh))
    rhs = np.vdot(orc.nfft_adjoint(N, n, m, x, fv), fh)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    lhs = np.vdot(fv, orc.ndft_trafo(N, x, fh))
    rhs = np.vdot(orc.ndft_adjoint(N, x, fv), fh)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_nfft_nonpow2_bandwidth(orc):
    N, m = [12, 20], 4
    n = [32, 64]
    r = rng(9)
    x = r.uniform(-0.5, 0.5, (200, 2))
    fh = crandn(r, 240)
    assert orc.rel_l2(orc.ndft_trafo(N, x, fh), orc.nfft_trafo(N, n, m, x, fh)) < 1e-4


@pytest.mark.parametrize("N,n,m,band", [([96], [192], 6, 1e-6), ([96, 64], [192, 128], 4, 3e-4),
                                        ([24, 16, 24], [48, 32, 72], 4, 3e-4)])
def test_nfft_on_grids_of_2a3b_cells(orc, N, n, m, band):
    """n = 2^a 3^b (sigma = 2 exactly for N = 96): the NFFT stays within the
    window band of the NDFT and keeps adjointness."""
    r = rng(17)
    x = r.uniform(-0.5, 0.5, (300, len(N)))
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, 300)
    assert orc.rel_l2(orc.ndft_trafo(N, x, fh), orc.nfft_trafo(N, n, m, x, fh)) < band
    assert orc.rel_l2(orc.ndft_adjoint(N, x, fv), orc.nfft_adjoint(N, n, m, x, fv)) < band
    lhs = np.vdot(fv, orc.nfft_trafo(N, n, m, x, fh))
    rhs = np.vdot(orc.nfft_adjoint(N, n, m, x, fv), fh)
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_edge_nodes(orc):
    # x = -0.5 exactly and both signed zeros (the radial trajectory's centre)
    N, m = [16, 16], 4
    n = orc.default_n(N)
    x = np.array([[-0.5, -0.5], [0.0, 0.0], [-0.0, 0.0], [0.0, -0.0], [-0.0, -0.0],
                  [0.49999999, -0.5]])
    fh = crandn(rng(10), 256)
    e = orc.rel_l2(orc.ndft_trafo(N, x, fh), orc.nfft_trafo(N, n, m, x, fh))
    assert e < 3e-4


# ----------------------------------------------------- planted mistakes ----
def _pipeline(orc, N, n, m, x, fh, sign=-1, rolloff=True, swap_axes=False):
    g = orc.rolloff_pad(N, n, m, fh)
    if not rolloff:
        big = np.zeros(int(np.prod(n)), complex)
        nz = g != 0
        big[nz] = orc.rolloff_pad(N, n, m, np.ones_like(fh))[nz]
        g[nz] = g[nz] / big[nz]
    g = orc.fft(n, g, sign)
    xx = x[:, ::-1].copy() if swap_axes else x
    return orc.gather(N, n, m, g, xx)


@pytest.mark.parametrize("mistake", ["sign", "rolloff", "swap_axes"])
def test_planted_mistakes_fail_the_band(orc, mistake):
    N, m = [16, 8], 4
    n = orc.default_n(N)
    r = rng(11)
    x = r.uniform(-0.5, 0.5, (300, 2))
    fh = crandn(r, 128)
    ref = orc.ndft_trafo(N, x, fh)
    good = _pipeline(orc, N, n, m, x, fh)
    assert orc.rel_l2(ref, good) < 3e-4
    bad = _pipeline(orc, N, n, m, x, fh, sign=+1 if mistake == "sign" else -1,
                    rolloff=mistake != "rolloff", swap_axes=mistake == "swap_axes")
    assert orc.rel_l2(ref, bad) > 1e-2


# ------------------------------------------- reference's recorded bands ----
def _bands(lo=0, hi=14):
    """Recorded rows with lo <= log2 N_total <= hi (one per direction, m, d, size);
    the larger ones run on the GPU (tests/test_gpu_reference_bands.py)."""
    with open(os.path.join(HERE, "golden", "reference_err_bands.json")) as fh:
        rows = json.load(fh)["rows"]
    seen, out = set(), []
    for r in rows:
        key = (r["direction"], r["m"], r["d"], r["log2_N_total"])
        L = r["log2_N_total"]
        if key in seen or L > hi or L < lo or L % r["d"]:
            continue
        seen.add(key)
        out.append(r)
    return out


@pytest.mark.parametrize("row", _bands(), ids=lambda r: r["source"])
def test_reference_recorded_error_band(orc, row):
    """Same drand48 inputs as the reference's driver; rel-l_inf NDFT vs NFFT lands
    in the band its FP64 run recorded (factor 3 for the adjoint rows; the forward
    driver that produced the rows differs from the shipped one, so factor 12).
    At m = 12 the recorded values are roundoff of the reference's FG_PSI
    recurrence, not window error, so only the upper bound applies there."""
    d, L, m = row["d"], row["log2_N_total"], row["m"]
    N = [2 ** (L // d)] * d
    M = 2 ** L
    n = orc.default_n(N)
    r = orc.Drand48(0)
    if row["direction"] == "forward":
        fh, x = r.example_data(N, M)
        e = orc.rel_linf(orc.ndft_trafo(N, x, fh), orc.nfft_trafo(N, n, m, x, fh))
        fac = 12.0
    else:
        f, x = r.example_data_ad(M, d)
        e = orc.rel_linf(orc.ndft_adjoint(N, x, f), orc.nfft_adjoint(N, n, m, x, f))
        fac = 3.0
    rec = row["err1_rel_linf"]
    assert e <= rec * fac
    if m <= 8:
        assert e >= rec / fac


# -------------------------------------------------------------- gradients --
def _away_from_jumps(n, x, h):
    # the window index set jumps where n x crosses a half-integer
    fr = (n * x) % 1.0
    return np.abs(fr - 0.5) > 4 * h * n


@pytest.mark.parametrize("N,m", [([32], 4), ([16, 16], 4), ([8, 8, 8], 3)])
def test_gather_dx_matches_finite_differences(orc, N, m):
    r = rng(12)
    d = len(N)
    n = orc.default_n(N)
    M = 40
    x = r.uniform(-0.45, 0.45, (M, d))
    fh = crandn(r, int(np.prod(N)))
    dfdx = orc.nfft_dx(N, n, m, x, fh)
    h = 1e-6
    for t in range(d):
        ok = _away_from_jumps(n[t], x[:, t], h)
        xp, xm = x.copy(), x.copy()
        xp[:, t] += h
        xm[:, t] -= h
        fd = (orc.nfft_trafo(N, n, m, xp, fh) - orc.nfft_trafo(N, n, m, xm, fh)) / (2 * h)
        err = np.abs(fd - dfdx[:, t])[ok]
        assert np.max(err) <= 1e-6 * np.max(np.abs(dfdx[:, t]))


def test_ndft_dx_matches_finite_differences(orc):
    r = rng(13)
    N = [8, 6]
    x = r.uniform(-0.5, 0.5, (20, 2))
    fh = crandn(r, 48)
    g = orc.ndft_dx(N, x, fh)
    h = 1e-6
    for t in range(2):
        xp, xm = x.copy(), x.copy()
        xp[:, t] += h
        xm[:, t] -= h
        fd = (orc.ndft_trafo(N, xp, fh) - orc.ndft_trafo(N, xm, fh)) / (2 * h)
        assert np.max(np.abs(fd - g[:, t])) <= 1e-7 * np.max(np.abs(g[:, t]))


def test_nfft_dx_near_ndft_dx(orc):
    r = rng(14)
    N, m = [32, 32], 6
    n = orc.default_n(N)
    x = r.uniform(-0.5, 0.5, (100, 2))
    fh = crandn(r, 1024)
    a = orc.nfft_dx(N, n, m, x, fh)
    b = orc.ndft_dx(N, x, fh)
    assert orc.rel_l2(b.reshape(-1), a.reshape(-1)) < 1e-4


# ------------------------------------------------------- Kaiser-Bessel ----
def test_bessel_i0_matches_scipy(orc):
    from scipy.special import i0

    for z in [0.0, 1e-3, 0.5, 3.0, 9.42, 18.85, 40.0, 75.4]:
        assert abs(orc.bessel_i0(z) / i0(z) - 1.0) <= 1e-14


def test_kaiser_bessel_window_shape(orc):
    m, b = 4, orc.kb_shape(64, 128)
    assert abs(b - 1.5 * np.pi) <= 1e-15
    # symmetric, continuous through |s| = m (value b / pi), series branch agrees
    for s in [0.3, 1.7, 3.9]:
        assert orc.phi_kb(s, m, b) == orc.phi_kb(-s, m, b)
    for eps in [1e-3, 1e-5, 1e-7, 1e-9]:
        # |phi'(m)| = b^3 m / (3 pi) ~ 44 here: the value moves linearly through b / pi
        assert abs(orc.phi_kb(m - eps, m, b) - b / np.pi) <= 50 * eps + 1e-12
        assert abs(orc.phi_kb(m + eps, m, b) - b / np.pi) <= 50 * eps + 1e-12
    # phi(0) closed form, derivative by central differences (both branches, and near m)
    assert abs(orc.phi_kb(0.0, m, b) - np.sinh(b * m) / (np.pi * m)) <= 1e-14 * np.sinh(b * m)
    for s in [0.0, 0.8, 2.5, 3.99, 3.99999, 4.00001, 4.3, 5.2]:
        h = 1e-6
        fd = (orc.phi_kb(s + h, m, b) - orc.phi_kb(s - h, m, b)) / (2 * h)
        an = orc.dphi_kb(s, m, b)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-6 * orc.phi_kb(0.0, m, b) * 1e-6


@pytest.mark.parametrize("d", [1, 2, 3])
def test_kaiser_bessel_nfft_converges_to_the_ndft(orc, d):
    """The Kaiser-Bessel NFFT reaches the NDFT far faster in m than the
    Gaussian (a wrong shape, roll-off or normalisation breaks this)."""
    N = {1: [64], 2: [32, 32], 3: [16, 16, 16]}[d]
    n = [2 * v for v in N]
    r = rng(30 + d)
    x = r.uniform(-0.5, 0.5, (300, d))
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, 300)
    nd, na = orc.ndft_trafo(N, x, fh), orc.ndft_adjoint(N, x, fv)
    bands = {2: 1e-3, 4: 3e-7, 6: 1e-10, 8: 1e-13}
    with orc.window("kaiser_bessel"):
        for m, band in bands.items():
            assert orc.rel_l2(nd, orc.nfft_trafo(N, n, m, x, fh)) <= band, m
            assert orc.rel_l2(na, orc.nfft_adjoint(N, n, m, x, fv)) <= band, m
        lhs = np.vdot(fv, orc.nfft_trafo(N, n, 4, x, fh))
        rhs = np.vdot(orc.nfft_adjoint(N, n, 4, x, fv), fh)
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # the Gaussian at the same m is far less accurate
    assert orc.rel_l2(nd, orc.nfft_trafo(N, n, 4, x, fh)) > 1e-5


def test_kaiser_bessel_gradient_matches_differences(orc):
    N, n, m = [16, 16], [32, 32], 4
    r = rng(40)
    x = r.uniform(-0.4, 0.4, (20, 2))
    fh = crandn(r, 256)
    with orc.window("kaiser_bessel"):
        dfdx = orc.nfft_dx(N, n, m, x, fh)
        h = 1e-7
        for j, t in [(0, 0), (3, 1), (7, 0), (11, 1)]:
            xp, xm = x.copy(), x.copy()
            xp[j, t] += h
            xm[j, t] -= h
            fd = (orc.nfft_trafo(N, n, m, xp, fh)[j] - orc.nfft_trafo(N, n, m, xm, fh)[j]) / (2 * h)
            assert abs(fd - dfdx[j, t]) <= 1e-6 * abs(dfdx[j, t]) + 1e-8


# ------------------------------------------------------------ sinc power --
def test_bspline_matches_scipy(orc):
    """The sinc window's Fourier transform, the centred cardinal B-spline,
    against scipy's B-spline basis element on the same unit knots."""
    from scipy.interpolate import BSpline

    for p in [2, 3, 4, 6, 8, 12, 16]:
        b = BSpline.basis_element(np.arange(p + 1) - p / 2, extrapolate=False)
        for x in np.linspace(-p / 2 + 1e-3, p / 2 - 1e-3, 41):
            assert abs(orc.bspline(p, x) - float(b(x))) <= 1e-14
        assert orc.bspline(p, p / 2 + 0.1) == 0.0 and orc.bspline(p, -p / 2) == 0.0


@pytest.mark.parametrize("m", [2, 3, 4])
def test_sinc_window_fourier_pair(orc, m):
    """phi(s) = a sinc(pi a s)^(2m) has Fourier transform M_2m(nu / a): checked
    by quadrature of the oracle's window (a wrong shape a, power or
    normalisation breaks it)."""
    a = orc.sinc_shape(64, 128, m)
    assert abs(a - 3.0 / (4.0 * m)) <= 1e-16
    S, h = 400.0 / a, 0.01 / a
    s = np.arange(-S, S + h / 2, h)
    phi = np.array([orc.phi_sinc(v, m, a) for v in s])
    for nu in [0.0, 0.05, 0.1, 0.2, 0.3, 0.37]:
        ft = np.sum(phi * np.cos(2 * np.pi * nu * s)) * h
        assert abs(ft - orc.bspline(2 * m, nu / a)) <= 1e-9


def test_sinc_window_derivative(orc):
    m, a = 4, orc.sinc_shape(64, 128, 4)
    for s in [0.0, 1e-3, 0.05, 0.4, 1.3, 2.9, 4.6, -3.2]:
        hh = 1e-6
        fd = (orc.phi_sinc(s + hh, m, a) - orc.phi_sinc(s - hh, m, a)) / (2 * hh)
        assert abs(fd - orc.dphi_sinc(s, m, a)) <= 1e-8
        assert orc.phi_sinc(s, m, a) == orc.phi_sinc(-s, m, a)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_sinc_nfft_converges_to_the_ndft(orc, d):
    """The sinc-power NFFT converges to the NDFT geometrically in m, and at
    m = 4 is an order of magnitude closer than the Gaussian."""
    N = {1: [64], 2: [32, 32], 3: [16, 16, 16]}[d]
    n = [2 * v for v in N]
    r = rng(50 + d)
    x = r.uniform(-0.5, 0.5, (300, d))
    fh = crandn(r, int(np.prod(N)))
    fv = crandn(r, 300)
    nd, na = orc.ndft_trafo(N, x, fh), orc.ndft_adjoint(N, x, fv)
    bands = {2: 6e-3, 3: 4e-4, 4: 2.5e-5, 6: 1e-7, 8: 4e-10}
    with orc.window("sinc"):
        errs = []
        for m, band in bands.items():
            ef = orc.rel_l2(nd, orc.nfft_trafo(N, n, m, x, fh))
            ea = orc.rel_l2(na, orc.nfft_adjoint(N, n, m, x, fv))
            assert ef <= band and ea <= band, (m, ef, ea)
            errs.append(ef)
        assert all(b < a / 5 for a, b in zip(errs, errs[1:]))
        lhs = np.vdot(fv, orc.nfft_trafo(N, n, 4, x, fh))
        rhs = np.vdot(orc.nfft_adjoint(N, n, 4, x, fv), fh)
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
        e_sinc = orc.rel_l2(nd, orc.nfft_trafo(N, n, 4, x, fh))
    assert orc.rel_l2(nd, orc.nfft_trafo(N, n, 4, x, fh)) > 5 * e_sinc


def test_sinc_gradient_matches_differences(orc):
    N, n, m = [16, 16], [32, 32], 4
    r = rng(60)
    x = r.uniform(-0.4, 0.4, (20, 2))
    fh = crandn(r, 256)
    with orc.window("sinc"):
        dfdx = orc.nfft_dx(N, n, m, x, fh)
        h = 1e-7
        for j, t in [(0, 0), (3, 1), (7, 0), (11, 1)]:
            xp, xm = x.copy(), x.copy()
            xp[j, t] += h
            xm[j, t] -= h
            fd = (orc.nfft_trafo(N, n, m, xp, fh)[j] - orc.nfft_trafo(N, n, m, xm, fh)[j]) / (2 * h)
            assert abs(fd - dfdx[j, t]) <= 1e-6 * abs(dfdx[j, t]) + 1e-8
