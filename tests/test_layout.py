"""Every plane-tiled LDS layout the library instantiates, checked on the CPU:
lane-table coverage, LDS bank rules, bounds of every access over all node
origins in the tile (gather, spread and held-unit spread), dummy units and
the slot-block partition (tests/layout/check_layout.cpp, compiled from
cunfft_amd/csrc/grid_tile.hpp for the host)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_every_instantiated_layout(tmp_path):
    exe = tmp_path / "check_layout"
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O1",
                    "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "cunfft_amd", "csrc"),
                    os.path.join(ROOT, "tests", "layout", "check_layout.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-6000:])
    assert r.returncode == 0 and r.stdout.strip().endswith("LAYOUT_OK")
    # the headline layouts (FP32 3D, W = 10, 8^3 tiles) are conflict-free
    for line in r.stdout.splitlines():
        if line.startswith("fp32 3D W=10 V=0"):
            assert "conflicts read 1 write 1" in line, line
