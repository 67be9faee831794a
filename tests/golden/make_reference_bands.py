"""Extract the reference's recorded error bands into reference_err_bands.json.

Reads the result tables the reference committed under Tests/Runtime/ (text
data, written by Tests/Scripts/simpleTest.cpp:395-433 as
" [m, TX x TY]  d  log2(N_total)  t_nfft t_cundft t_cunfft | err1 err2"),
keeps rows whose err1 (relative l_inf error CUNDFT vs CUNFFT, FP64,
src/cuda_util.cpp:237-258) is a valid number (< 0.5; err1 = 1.0 marks a
failed run), and writes them as a small JSON fixture.  Each row records the
file and line it came from.  Run in a checkout that has the reference at
/root/reference:  python tests/golden/make_reference_bands.py
"""
import json
import os
import re

REF = "/root/reference/Tests/Runtime"
FILES = {  # file -> direction
    "test_Runtime_m6.txt": "forward",
    "test_m6_ad.txt": "adjoint",
    "test_Runtime_m12.txt": "forward",
    "test_m12_ad.txt": "adjoint",
}
ROW = re.compile(r"^\s*\[(\d+),\s*\d+\s*x\s*\d+\]\s+(\d)\s+(\d+)\s+\S+\s+\S+\s+\S+\s*\|\s*(\S+)\s+(\S+)")


def main():
    rows = []
    for fn, direction in FILES.items():
        with open(os.path.join(REF, fn)) as fh:
            for lineno, line in enumerate(fh, 1):
                mm = ROW.match(line)
                if not mm:
                    continue
                m, d, logn, e1, e2 = mm.groups()
                e1 = float(e1)
                if not e1 < 0.5:
                    continue
                rows.append({"source": f"Tests/Runtime/{fn}:{lineno}", "direction": direction,
                             "m": int(m), "d": int(d), "log2_N_total": int(logn),
                             "err1_rel_linf": e1, "err2_rel_linf": float(e2)})
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_err_bands.json")
    with open(out, "w") as fh:
        json.dump({"note": "rows of the reference's recorded FP64 CUNDFT-vs-CUNFFT errors; "
                           "inputs are glibc drand48 from state 0 (getExampleData_uniDistr), "
                           "M = N_total, n = 2*next_pow2(N)", "rows": rows}, fh, indent=1)
    print(f"wrote {len(rows)} rows to {out}")


if __name__ == "__main__":
    main()
