"""The window-cell rule of the gridding kernels, run on the CPU.

``window_cell`` (cunfft_amd/csrc/device_common.hpp) is the one place where a
node coordinate becomes a grid index: the node binning and the untiled
kernels all call it.  It is compiled here for the host (hipcc, host pass
only) and checked on the CPU, before any GPU run:

* every result lies in [0, n), for any double -- nodes far outside
  [-1/2, 1/2), +-0, +-inf, NaN, huge values;
* for finite nodes it is the reference's rule, the oracle's
  u = llrint(n x) - signbit(x) - m (src/cunfft_kernel.cu:175-182) taken mod n
  (the reference's mask with n - 1, src/cunfft_kernel.cu:528-531), with the
  fraction c = n x - u.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROGRAM = r"""
#include "device_common.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>
// argv: n m infile outfile; in: doubles; out: (int64 cell, double fraction) per node
int main(int argc, char **argv)
{
    const long long n = atoll(argv[1]);
    const int m = atoi(argv[2]);
    FILE *fi = fopen(argv[3], "rb");
    std::vector<double> x;
    double v;
    while (fread(&v, sizeof v, 1, fi) == 1) x.push_back(v);
    fclose(fi);
    FILE *fo = fopen(argv[4], "wb");
    for (double xv : x) {
        double c;
        const long long w = (long long)cunfft::window_cell(xv, n, m, c);
        fwrite(&w, sizeof w, 1, fo);
        fwrite(&c, sizeof c, 1, fo);
    }
    fclose(fo);
    return 0;
}
"""


@pytest.fixture(scope="module")
def window_cell_host(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.fail("hipcc is needed to compile the host replica")
    d = tmp_path_factory.mktemp("wc")
    src = d / "wc.cpp"
    src.write_text(PROGRAM)
    exe = d / "wc"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "cunfft_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)

    def run(n, m, x):
        fi, fo = d / "in.bin", d / "out.bin"
        np.ascontiguousarray(x, dtype=np.float64).tofile(fi)
        subprocess.run([str(exe), str(n), str(m), str(fi), str(fo)], check=True)
        rec = np.fromfile(fo, dtype=np.dtype([("w", "<i8"), ("c", "<f8")]))
        return rec["w"], rec["c"]

    return run


def special_nodes():
    big = [1e300, -1e300, 2.0 ** 60, -(2.0 ** 60), 1e15, -1e15, 2.0 ** 40 + 0.5]
    return np.array([0.0, -0.0, -0.5, 0.5, 0.25, -0.25, np.inf, -np.inf, np.nan, -np.nan,
                     5e-324, -5e-324] + big)


@pytest.mark.parametrize("n,m", [(256, 4), (2048, 6), (192, 4), (96, 1), (64, 8), (8388608, 4)])
def test_cell_is_in_the_grid_for_any_node(window_cell_host, n, m):
    r = np.random.default_rng(n + m)
    x = np.concatenate([special_nodes(), r.uniform(-1.5, 1.5, 4000),
                        r.uniform(-1e6, 1e6, 1000), (np.arange(-40, 41) + 0.5) / n,
                        np.arange(-40, 41) / n])
    w, _ = window_cell_host(n, m, x)
    assert w.min() >= 0 and w.max() < n


@pytest.mark.parametrize("n,m", [(256, 4), (2048, 6), (192, 4), (96, 2)])
def test_cell_is_the_reference_rule_mod_n(window_cell_host, orc, n, m):
    r = np.random.default_rng(7 * n + m)
    # nodes inside and outside [-1/2, 1/2), half-integer n x (round half to
    # even), signed zeros
    x = np.concatenate([r.uniform(-1.5, 1.5, 3000), (np.arange(-3 * n // 2, 3 * n // 2) + 0.5) / n,
                        np.array([0.0, -0.0, -0.5, -1.5, 1.25])])
    w, c = window_cell_host(n, m, x)
    u = np.array([orc.window_start(n, float(v), m) for v in x], dtype=np.int64)
    assert np.array_equal(w, np.mod(u, n))
    assert np.array_equal(c, n * x - u.astype(np.float64))


def test_non_finite_nodes_keep_non_finite_fractions(window_cell_host):
    x = np.array([np.nan, np.inf, -np.inf])
    w, c = window_cell_host(256, 4, x)
    assert np.all(~np.isfinite(c))
    assert np.all((w >= 0) & (w < 256))
