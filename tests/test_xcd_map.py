"""The XCD-aware workgroup -> chunk map of the tile-staging kernels, on the CPU.

``xcd_chunk`` (cunfft_amd/csrc/bin.hpp) deals the n chunks to the 8 XCDs
(workgroup b runs on XCD b mod 8) in groups of kXcdGroup = 128 consecutive
chunks, group g to XCD g mod 8.  The kernels that use it (the plane-tiled
gather and x-gradient) launch ``xcd_grid(max_chunks)`` workgroups for a chunk
count n <= max_chunks known only on the device.  Checked here, compiled for
the host (hipcc, host pass only), for chunk counts around the group and grid
boundaries and for large counts, with max_chunks from n to n + 20:

* the active workgroups (chunk >= 0) cover every chunk in [0, n) exactly
  once -- no chunk lost, none done twice -- and every chunk index is < n;
* every chunk of group g runs on XCD g mod 8, and the workgroups of one XCD
  take their chunks in increasing order (tiles in order);
* the XCDs' shares differ by at most one group.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROGRAM = r"""
#include "bin.hpp"
#include <cstdio>
#include <cstdlib>
// argv: n max_chunks; prints one chunk index (or -1) per workgroup of the grid
int main(int argc, char **argv)
{
    const unsigned n = (unsigned)atoll(argv[1]);
    const long long mx = atoll(argv[2]);
    const unsigned grid = cunfft::xcd_grid(mx);
    for (unsigned b = 0; b < grid; ++b) printf("%lld\n", (long long)cunfft::xcd_chunk(b, n));
    return 0;
}
"""


@pytest.fixture(scope="module")
def xcd_map(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.fail("hipcc is needed to compile the host replica")
    d = tmp_path_factory.mktemp("xcd")
    src = d / "xcd.cpp"
    src.write_text(PROGRAM)
    exe = d / "xcd"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2",
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "cunfft_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)

    def run(n, mx):
        out = subprocess.run([str(exe), str(n), str(mx)], check=True, capture_output=True, text=True)
        return np.array(out.stdout.split(), dtype=np.int64)

    return run


def check(c, n):
    grid = len(c)
    assert grid % 8 == 0
    act = c[c >= 0]
    assert np.all(act < n)
    assert np.array_equal(np.sort(act), np.arange(n)), n
    shares = []
    for x in range(8):
        mine = c[x::8]
        mine = mine[mine >= 0]
        assert np.all((mine // 128) % 8 == x), (n, x)
        assert np.all(np.diff(mine) > 0), (n, x)
        shares.append(len(mine))
    assert max(shares) - min(shares) <= 128


def test_every_chunk_once_small_counts(xcd_map):
    counts = sorted(set(range(0, 40)) | {127, 128, 129, 1023, 1024, 1025, 1151, 1152, 1153, 2047, 3000})
    for n in counts:
        for mx in (n, n + 1, n + 7, n + 8, n + 20):
            check(xcd_map(n, mx), n)


def test_every_chunk_once_large_counts(xcd_map):
    for n, mx in ((36111, 40960), (40960, 40960), (40953, 40960), (1, 40960)):
        check(xcd_map(n, mx), n)
