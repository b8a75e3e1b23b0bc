"""Write tests/golden/bessel_ref_sample.npz: every 250th row of ABCMB's
tabulated phi0/phi1/phi2 (abcmb/bessel_tab/phi*.txt, 161 multipoles) plus the
multipole list.  These are the reference's own data values; the test
tests/test_bessel.py checks the scipy-regenerated tables against them.

Usage: python make_bessel_golden.py /path/to/reference/abcmb/bessel_tab
"""
import os
import sys

import numpy as np

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/abcmb/bessel_tab"
rows = np.arange(0, 5000, 250)
out = {"l": np.loadtxt(os.path.join(src, "l.txt")).astype(np.int32), "rows": rows}
for f in range(3):
    out[f"phi{f}"] = np.loadtxt(os.path.join(src, f"phi{f}.txt"))[rows]
np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "bessel_ref_sample.npz"), **out)
print("ok")
