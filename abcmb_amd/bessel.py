"""Tabulated spherical Bessel functions for the line-of-sight integral.

ABCMB ships phi0 = j_l, phi1 = j_l' and phi2 = (3 j_l'' + j_l)/2 on 5000
uniform x-points per sampled multipole (abcmb/bessel_tab/*.txt, loaded at
abcmb/spectrum.py:20-27).  This module regenerates the same tables from their
grid definition (``data/bessel_ranges.json``: the multipole list and the
per-multipole x-ranges) with scipy, and caches them in
``data/bessel_tables.npz``.  Layout for the GPU: ``tab[3][n_l][nx]`` and
``rng[3][n_l][2]``.
"""
from __future__ import annotations

import json
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
RANGES = os.path.join(_DATA, "bessel_ranges.json")
CACHE = os.path.join(_DATA, "bessel_tables.npz")


def _phi(which: int, l: int, x: np.ndarray) -> np.ndarray:
    from scipy.special import spherical_jn
    if which == 0:
        return spherical_jn(l, x)
    if which == 1:
        return spherical_jn(l, x, derivative=True)
    j = spherical_jn(l, x)
    jp = spherical_jn(l, x, derivative=True)
    # j'' from the spherical Bessel equation x^2 j'' + 2x j' + (x^2 - l(l+1)) j = 0
    safe = np.where(x > 0, x, 1.0)
    jpp = -(2 * safe * jp + (safe ** 2 - l * (l + 1)) * j) / safe ** 2
    # limit at x = 0: j_l''(0) = 2/15 for l = 2, else 0 (l >= 2 here)
    jpp = np.where(x > 0, jpp, 2.0 / 15.0 if l == 2 else 0.0)
    return (3 * jpp + j) / 2


def build_tables(write_cache: bool = True):
    with open(RANGES) as fh:
        spec = json.load(fh)
    ls = np.asarray(spec["l"], dtype=np.int32)
    nx = int(spec["nx"])
    tab = np.zeros((3, ls.size, nx))
    rng = np.zeros((3, ls.size, 2))
    for f in range(3):
        for i, l in enumerate(ls):
            lo, hi = spec[f"range{f}"][i]
            rng[f, i] = (lo, hi)
            tab[f, i] = _phi(f, int(l), np.linspace(lo, hi, nx))
    if write_cache:
        np.savez(CACHE, l=ls, tab=tab, rng=rng)
    return ls, tab, rng


def load_tables():
    """(l [n_l] int32, tab [3][n_l][nx], rng [3][n_l][2])."""
    if os.path.exists(CACHE):
        d = np.load(CACHE)
        return d["l"], d["tab"], d["rng"]
    return build_tables(write_cache=True)
