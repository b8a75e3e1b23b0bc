"""Extract the multipole list and x-ranges of ABCMB's tabulated spherical Bessel
functions (abcmb/bessel_tab/{l,xphi0,xphi1,xphi2}.txt, loaded at
abcmb/spectrum.py:20-27) into abcmb_amd/data/bessel_ranges.json, the grid
definition the package ships.  The table values themselves are regenerated
with scipy by abcmb_amd.bessel; only the grid definition is taken from the
reference data files.  Run once by hand; nothing in the package imports it.

Usage: python make_bessel_ranges.py /path/to/reference/abcmb/bessel_tab
"""
import json
import os
import sys

import numpy as np

src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/abcmb/bessel_tab"
ls = np.loadtxt(os.path.join(src, "l.txt")).astype(int)
out = {"l": ls.tolist(), "nx": 5000}
for f in range(3):
    x = np.loadtxt(os.path.join(src, f"xphi{f}.txt"))
    assert x.shape == (5000, ls.size)
    out[f"range{f}"] = [[float(x[:, i].min()), float(x[:, i].max())] for i in range(ls.size)]
root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
with open(os.path.join(root, "abcmb_amd", "data", "bessel_ranges.json"), "w") as fh:
    json.dump(out, fh)
print("wrote", ls.size, "multipoles")
