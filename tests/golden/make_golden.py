"""Write tests/golden/golden_cases.npz: small seeded NFFT/NDFT input-output vectors.

Outputs come from the pinned CPU oracle (oracle/, FP64): the direct NDFT and
the NFFT restatement, forward and adjoint.  The GPU tests compare against
these files, so the GPU box needs neither the reference nor a long CPU run.
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as orc  # noqa: E402

CASES = {
    "d1_m4": dict(N=[32], n=[64], m=4, M=64, seed=101),
    "d1_m6": dict(N=[64], n=[128], m=6, M=96, seed=102),
    "d2_m4": dict(N=[16, 16], n=[32, 32], m=4, M=200, seed=103),
    "d2_nonsq_m3": dict(N=[12, 20], n=[32, 64], m=3, M=150, seed=104),
    "d3_m4": dict(N=[8, 8, 8], n=[16, 16, 16], m=4, M=300, seed=105),
    "d3_m2": dict(N=[8, 6, 4], n=[16, 16, 8], m=2, M=100, seed=106),
}


def main():
    out = {}
    for name, c in CASES.items():
        r = np.random.default_rng(c["seed"])
        d = len(c["N"])
        Nt = int(np.prod(c["N"]))
        x = r.uniform(-0.5, 0.5, (c["M"], d))
        x[0] = -0.5          # edge of the torus
        if c["M"] > 3:
            x[1] = 0.0       # +0.0 and -0.0 use different window index sets
            x[2] = -0.0
        fh = r.standard_normal(Nt) + 1j * r.standard_normal(Nt)
        fv = r.standard_normal(c["M"]) + 1j * r.standard_normal(c["M"])
        out[f"{name}/N"] = np.array(c["N"])
        out[f"{name}/n"] = np.array(c["n"])
        out[f"{name}/m"] = np.array(c["m"])
        out[f"{name}/x"] = x
        out[f"{name}/fhat"] = fh
        out[f"{name}/f"] = fv
        out[f"{name}/ndft_trafo"] = orc.ndft_trafo(c["N"], x, fh)
        out[f"{name}/ndft_adjoint"] = orc.ndft_adjoint(c["N"], x, fv)
        out[f"{name}/nfft_trafo"] = orc.nfft_trafo(c["N"], c["n"], c["m"], x, fh)
        out[f"{name}/nfft_adjoint"] = orc.nfft_adjoint(c["N"], c["n"], c["m"], x, fv)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_cases.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(CASES)} cases)")


if __name__ == "__main__":
    main()
