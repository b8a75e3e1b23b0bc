"""HyRex recombination restated in numpy/scipy -- TEST INFRASTRUCTURE ONLY.

The checker for the host producer's recombination (``abcmb_amd/csrc/hyrex.cpp``):
the same stages of the reference (abcmb/hyrex/helium.py, hydrogen.py,
recomb_functions.py, hyrex.py) written independently, with other numerics:
the rate equations integrated by ``scipy.integrate.solve_ivp`` (Radau,
rtol 1e-11, dense output read at the grid points), the table interpolation by
``scipy.ndimage.map_coordinates`` (order 1, the routine family the reference
calls), and the post-Saha derivatives by central differences.  Each stage
stops at the first grid point where its event is true, as the product does.
"""
from __future__ import annotations

import math
import os

import numpy as np
from scipy.integrate import solve_ivp
from scipy.ndimage import map_coordinates

from abcmb_amd import constants as K

_DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "abcmb_amd", "data", "hyrex")
DX = 5.0e-4
X0 = -math.log(8001.0)
N_GRID = np.arange(X0, 0.0, DX).size
TR_MIN, TR_MAX, NTR = 0.004, 0.4, 100
TM_MIN, TM_MAX, NTM = 0.1, 1.0, 40

_alpha = np.loadtxt(os.path.join(_DATA, "Alpha_inf.dat")).reshape(NTR, NTM, 4)
_lnA2s, _lnA2p = np.log(_alpha[:, :, 0]), np.log(_alpha[:, :, 1])
_lnR = np.log(np.loadtxt(os.path.join(_DATA, "R_inf.dat")))
_swift = np.loadtxt(os.path.join(_DATA, "fit_swift.dat"))
_TR_axis = np.linspace(math.log(TR_MIN), math.log(TR_MAX), NTR)
_TM_axis = np.linspace(TM_MIN, TM_MAX, NTM)
LAM_LYA = 2 * math.pi * K.hbar * K.c / K.rydberg * 4 / 3
LYA_FREQ = K.rydberg * 3 / 4 / (2 * math.pi * K.hbar)


class Inputs:
    """T_CMB, n_H and H on the grid, linear interpolation (hyrex.py:16-90)."""

    def __init__(self, T, nh, H):
        self.T, self.nh, self.H = (np.asarray(a, dtype=np.float64) for a in (T, nh, H))
        self.x1 = X0 + (N_GRID - 1) * DX

    def _at(self, f, x):
        i = (x - X0) / (self.x1 - X0) * (N_GRID - 1)
        i = min(max(i, 1e-6), N_GRID - 1 - 1e-6)
        il = int(math.floor(i))
        iu = min(il + 1, N_GRID - 1)
        w = i - il
        return (1 - w) * f[il] + w * f[iu]

    def __call__(self, x):
        return self._at(self.T, x), self._at(self.nh, x), self._at(self.H, x)


def alpha_H(Tm):
    t = Tm / K.kB * 1e-4
    return 4.309e-13 * t ** -0.6166 / (1 + 0.6703 * t ** 0.53)


def beta_H(T):
    lam = 2 * math.pi * K.hbar * K.c / math.sqrt(2 * math.pi * K.mu_e * T)
    return lam ** -3 * math.exp(-K.rydberg / 4 / T) * alpha_H(T) / 4


def gamma_compton(xe, T, YHe):
    FHe = YHe * K.mH / K.mHe / (1 - YHe)
    return xe / (1 + FHe + xe) * 8 * K.thomson_xsec * 4 * K.stef_bolt * T ** 4 / (3 * K.me)


def _idx(x, axis):
    return float(np.interp(x, axis, np.arange(axis.size)))


def rates(T, Tm):
    """A2s, A2p, B2s, B2p, R2p2s, R2s2p (recomb_functions.py:180-258)."""
    iT, ir = _idx(math.log(T), _TR_axis), _idx(Tm / T, _TM_axis)
    mc = lambda t, j: float(np.exp(map_coordinates(t, [[iT], [j]], order=1, mode="constant", cval=0.0))[0])
    A2s, A2p = mc(_lnA2s, ir), mc(_lnA2p, ir)
    q = (2 * math.pi * K.mu_e * T / (2 * math.pi * K.hbar) ** 2) ** 1.5 / K.c ** 3
    e2 = math.exp(-K.rydberg / 4 / T)
    B2s = e2 * q * mc(_lnA2s, NTM - 1)
    B2p = e2 * q * mc(_lnA2p, NTM - 1) / 3
    R2p2s = math.exp(np.interp(math.log(T), _TR_axis, _lnR))
    return A2s, A2p, B2s, B2p, R2p2s, 3 * R2p2s


def dxe_emla(xe, T, Tm, H, nh, Delta):
    """hydrogen.py:497-536 with the steady state of :577-636."""
    A2s, A2p, B2s, B2p, R2p2s, R2s2p = rates(T, Tm)
    x1s = 1 - xe
    RLya = 8 * math.pi * H / 3 / nh / x1s / LAM_LYA ** 3
    R2p1s = RLya / (1 + Delta)
    b = math.exp(-K.E21 / T)
    M = np.array([[B2s + R2s2p + K.R2s1s, -R2p2s], [-R2s2p, B2p + R2p2s + R2p1s]])
    S = np.array([xe ** 2 * nh * A2s + x1s * b * K.R2s1s, xe ** 2 * nh * A2p + x1s * 3 * b * R2p1s])
    x2s, x2p = np.linalg.solve(M, S)
    return (x2s * B2s + x2p * B2p - xe ** 2 * nh * (A2s + A2p)) / H


def swift_delta(T, om_b, om_cdm, YHe, neff):
    s = _swift
    T_ax = K.kB * s[:, 0]
    g = lambda c: np.interp(T, T_ax, s[:, c])
    return (g(1) + (om_b + om_cdm - 0.14175) * g(2) + (om_b * (1 - YHe) - 0.01689) * g(3)
            + (neff - 3.046) * g(4))


def _fhe(YHe):
    return YHe / (1 - YHe) / 3.97153


def xHeII_saha(inp, YHe, lna):
    T, nh, _ = inp(lna)
    Tk = T / K.kB
    s = 4 * 2.414194e15 * Tk * math.sqrt(Tk) * math.exp(-285325.0 / Tk) / nh
    f = _fhe(YHe)
    return 2 * s * f / (1 + s) / (1 + math.sqrt(1 + 4 * s * f / (1 + s) ** 2))


def xH1_saha(inp, YHe, lna):
    T, nh, _ = inp(lna)
    xHeII = xHeII_saha(inp, YHe, lna)
    Tk = T / K.kB
    s = 2.4127161187130e15 * Tk * math.sqrt(Tk) * math.exp(-157801.37882 / Tk) / nh
    if s > 1e5:
        return (1 + xHeII) / s - (xHeII ** 2 + 3 * xHeII + 2) / s ** 2
    if s == 0:
        return 1.0
    u = 1 + xHeII / s
    return 1 - 2 / (u + math.sqrt(u * u + 4 / s))


def he_rate(inp, YHe, xe, lna):
    """helium.py:375-453."""
    f = _fhe(YHe)
    T, nh, H = inp(lna)
    xH1 = xH1_saha(inp, YHe, lna)
    xHeII = xe - (1 - xH1)
    Tk = T / K.kB
    s0 = 2.414194e15 * Tk * math.sqrt(Tk) / nh * 4
    s = s0 * math.exp(-285325.0 / Tk)
    y2s = math.exp(46090.0 / Tk) / s0
    y2p = math.exp(39101.0 / Tk) / s0 * 3
    etacinv = 9.15776e22 * H / (nh * xH1)
    lines = ((6.03e6, 19754.0), (1.06e8, 21539.0), (2.18e6, 28496.0), (3.37e7, 29224.0), (1.04e6, 32414.0),
             (1.51e7, 32781.0))
    g2pinc = 1.976e6 / (1 - math.exp(-6989.0 / Tk)) + sum(a / (math.exp(e / Tk) - 1) for a, e in lines)
    tau2p = 4.277e-8 * nh / H * (f - xHeII)
    tauc = g2pinc * tau2p / (4 * math.pi ** 2) / etacinv
    if 1 + math.pi ** 2 * tauc < 0:   # a trial state past x_HeII = f_He (the solver steps back)
        return float("nan")
    enh = math.sqrt(1 + math.pi ** 2 * tauc) + 7.74 * tauc / (1 + 70 * tauc)
    pesc = enh / tau2p + (1 - math.exp(-1.023e-7 * tau2p)) * (
        0.964525 * math.exp(2947.0 / Tk) - enh * math.exp(-6.14e13 / etacinv)) / tau2p
    ydown = 50.94 * y2s + 1.7989e9 * y2p * pesc
    return ydown * ((f - xHeII) * s - xHeII * (xHeII + 1 - xH1)) / H


def _deriv(fn, x, h):
    """Central difference with step h."""
    return (fn(x + h) - fn(x - h)) / (2 * h)


def _deriv_back(fn, x, h):
    """Second-order backward difference (the function ends just above x)."""
    return (3 * fn(x) - 4 * fn(x - h) + fn(x - 2 * h)) / (2 * h)


def _ode(rhs, t0, y0, n_max, stop):
    """Saved points t0 + k DX (k = 1..n_max) of the solution, until stop(t, y) at one."""
    ts = t0 + DX * np.arange(1, n_max + 1)
    sol = solve_ivp(rhs, (t0, ts[-1]), np.atleast_1d(y0), method="Radau", rtol=1e-11, atol=1e-15,
                    dense_output=True)
    out = []
    for t in ts:
        y = sol.sol(t)
        out.append(y)
        if stop(t, y):
            break
    return np.array(out)


def history(T, nh, H, YHe, omega_b, omega_cdm, neff):
    """x_e on the grid from its start and T_m from the EMLA-with-T_m stage:
    (xe [n], tm [m], first grid index of T_m, stage ends)."""
    inp = Inputs(T, nh, H)
    g = lambda j: X0 + j * DX
    xe, ends = [], []
    f = _fhe(YHe)
    n4 = int(np.sum(np.arange(X0, 0.0, DX) <= -math.log(3601.0)))
    for i in range(n4):                                             # HeIII -> HeII Saha
        Tc, n_h, _ = inp(g(i))
        Tk = Tc / K.kB
        s = 2.414194e15 * Tk * math.sqrt(Tk) * math.exp(-631462.7 / Tk) / n_h
        x3 = 2 * s * f / (1 + s + f) / (1 + math.sqrt(1 + 4 * s * f / (1 + s + f) ** 2))
        xe.append(1 + f + x3)
        if x3 < 1e-9:
            break
    ends.append(len(xe))
    for _ in range(4000):                                           # post-Saha HeII
        lna = g(len(xe))
        xH1, xHeII = xH1_saha(inp, YHe, lna), xHeII_saha(inp, YHe, lna)
        xs = 1 - xH1 + xHeII
        # he_rate ends where x_HeII reaches f_He (its optical depth vanishes)
        hx = 1e-3 * (f - xHeII) if f - xHeII > 0 else 1e-9
        x = xs + _deriv(lambda l: xHeII_saha(inp, YHe, l), lna, 1e-6) / _deriv_back(
            lambda v: he_rate(inp, YHe, v, lna), xs, hx)
        xe.append(x)
        if abs(xs - x) > 1e-5:
            break
    ends.append(len(xe))
    ys = _ode(lambda t, y: [he_rate(inp, YHe, y[0] + xH1_saha(inp, YHe, t), t)], g(len(xe) - 1), xe[-1], 4096,
              lambda t, y: y[0] - (1 - xH1_saha(inp, YHe, t)) < 1e-4)
    xe += list(ys[:, 0])
    ends.append(len(xe))
    for _ in range(800):                                            # post-Saha hydrogen
        lna = g(len(xe))
        Tc, n_h, Hc = inp(lna)
        ge = (2 * math.pi * K.mu_e * Tc) ** 1.5 / (2 * math.pi * K.hbar * K.c) ** 3 / n_h
        s = ge * math.exp(-K.rydberg / Tc)
        xs = (math.sqrt(s * s + 4 * s) - s) / 2
        dsaha = -(K.rydberg / Tc - 1.5) * xs ** 2 / (2 * xs + s)
        x = xs + dsaha / _deriv(lambda v: dxe_emla(v, Tc, Tc, Hc, n_h, 0.0), xs, 1e-4 * (1 - xs))
        xe.append(x)
        if abs(xs - x) > 1e-5:
            break
    ends.append(len(xe))

    def rhs_2g(t, y):
        Tc, n_h, Hc = inp(t)
        Tm = Tc * (1 - Hc / gamma_compton(y[0], Tc, YHe))
        return [dxe_emla(y[0], Tc, Tm, Hc, n_h, swift_delta(Tc, omega_b, omega_cdm, YHe, neff))]
    ys = _ode(rhs_2g, g(len(xe) - 1), xe[-1], 2048, lambda t, y: t > -math.log(701.0))
    xe += list(ys[:, 0])
    ends.append(len(xe))
    jt0 = len(xe)
    t0 = g(len(xe) - 1)
    Tc, _, Hc = inp(t0)
    y0 = [xe[-1], Tc * (1 - Hc / gamma_compton(xe[-1], Tc, YHe))]

    def rhs_emla(t, y):
        Tc, n_h, Hc = inp(t)
        return [dxe_emla(y[0], Tc, y[1], Hc, n_h, 0.0), (-2 * Hc * y[1] + gamma_compton(y[0], Tc, YHe) * (Tc - y[1])) / Hc]

    def stop_emla(t, y):
        Tc = inp(t)[0]
        return Tc < TR_MIN or min(y[1] / Tc, Tc / y[1]) < TM_MIN
    ys = _ode(rhs_emla, t0, y0, 8192, stop_emla)
    xe += list(ys[:, 0])
    tm = list(ys[:, 1])
    ends.append(len(xe))

    def rhs_tla(t, y):
        Tc, n_h, Hc = inp(t)
        rate = 8 * math.pi * Hc / (3 * n_h * (K.c / LYA_FREQ) ** 3)
        exc = 3 * rate / 4 + (1 - y[0]) * K.R2s1s / 4
        ion = (1 - y[0]) * beta_H(Tc)
        C = exc / (exc + ion)
        return [C * (beta_H(y[1]) * (1 - y[0]) - alpha_H(y[1]) * n_h * y[0] ** 2) / Hc,
                -2 * y[1] + gamma_compton(y[0], Tc, YHe) / Hc * (Tc - y[1])]
    ys = _ode(rhs_tla, g(len(xe) - 1), [xe[-1], tm[-1]], 8192, lambda t, y: t > 0.0)
    xe += list(ys[:, 0])
    tm += list(ys[:, 1])
    ends.append(len(xe))
    return np.array(xe), np.array(tm), jt0, ends
