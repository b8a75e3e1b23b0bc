"""Not-a-knot cubic spline as a linear operator.

ABCMB interpolates the sampled C_l to every integer multipole with an
interpax ``CubicSpline`` (abcmb/spectrum.py:598-601; not-a-knot end
conditions, extrapolating with the end cubics).  The interpolant is linear
in the data, so for fixed knots and query points it is a dense matrix that
the GPU applies with ``abcmb_apply_operator``.  The matrix is built here
once per model from the second-derivative (M) formulation.
"""
from __future__ import annotations

import numpy as np


def _second_derivs(x: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """M (n x m) of the not-a-knot spline through (x, Y[:, j]), n >= 4."""
    n = x.size
    h = np.diff(x)
    m = n - 2
    A = np.zeros((m, m))
    d = np.zeros((m,) + Y.shape[1:])
    for r in range(m):
        i = r + 1
        if r > 0:
            A[r, r - 1] = h[i - 1]
        A[r, r] = 2 * (h[i - 1] + h[i])
        if r < m - 1:
            A[r, r + 1] = h[i]
        d[r] = 6 * ((Y[i + 1] - Y[i]) / h[i] - (Y[i] - Y[i - 1]) / h[i - 1])
    # not-a-knot: M0 = M1 (1 + h0/h1) - (h0/h1) M2, and symmetrically at the end
    A[0, 0] = h[0] * (1 + h[0] / h[1]) + 2 * (h[0] + h[1])
    A[0, 1] = h[1] - h[0] ** 2 / h[1]
    h2, h3 = h[n - 2], h[n - 3]
    A[m - 1, m - 2] = h3 - h2 ** 2 / h3
    A[m - 1, m - 1] = 2 * (h3 + h2) + h2 * (1 + h2 / h3)
    Mi = np.linalg.solve(A, d)
    M = np.zeros((n,) + Y.shape[1:])
    M[1:-1] = Mi
    M[0] = M[1] * (1 + h[0] / h[1]) - h[0] / h[1] * M[2]
    M[-1] = M[-2] * (1 + h[-1] / h[-2]) - h[-1] / h[-2] * M[-3]
    return M


def evaluate(x: np.ndarray, Y: np.ndarray, xq: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    M = _second_derivs(x, Y)
    j = np.clip(np.searchsorted(x, xq, side="right") - 1, 0, x.size - 2)
    h = x[j + 1] - x[j]
    a = (x[j + 1] - xq) / h
    b = (xq - x[j]) / h
    sh = (-1,) + (1,) * (Y.ndim - 1)
    a = a.reshape(sh); b = b.reshape(sh); h = h.reshape(sh)
    return a * Y[j] + b * Y[j + 1] + ((a ** 3 - a) * M[j] + (b ** 3 - b) * M[j + 1]) * h ** 2 / 6.0


def operator(x: np.ndarray, xq: np.ndarray) -> np.ndarray:
    """Dense [len(xq), len(x)] matrix of the not-a-knot spline."""
    return evaluate(x, np.eye(len(x)), np.asarray(xq, dtype=np.float64))
