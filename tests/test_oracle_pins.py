"""CPU pins of the oracle (oracle/abcmb_oracle.c) against mathematics, scipy
and the reference's own data.  Nothing here needs a GPU.

What each test pins (a plausible mistake it would catch):
  - Kvaerno5 tableau: all order-5 conditions for b, order-4 for b_hat
    (a mistyped coefficient);
  - not-a-knot cubic spline: scipy.interpolate.CubicSpline (wrong end
    condition, wrong interval);
  - Wigner d: Jacobi-polynomial closed form (wrong recursion coefficient,
    sign convention, normalisation);
  - Bessel tables: the reference's tabulated values (abcmb/bessel_tab);
  - LOS lookups and k-quadrature: scipy.special.spherical_jn (wrong branch,
    E-mode factor, trapezoid weights);
  - perturbation RHS: the synchronous-gauge Einstein equations that the
    evolution does not use (Ma & Bertschinger 1995 eqs. 21c, 21d; a dropped
    term or wrong coefficient in any fluid equation up to l = 2);
  - ODE integrator: scipy Radau at tight tolerance (wrong stage, error
    estimate, dense output);
  - lensing: zero lensing potential returns the unlensed spectra and BB = 0
    (Gauss-Legendre/Wigner orthogonality).
"""
import math

import numpy as np
import pytest
from scipy.interpolate import CubicSpline
from scipy.special import eval_jacobi, spherical_jn, gammaln

from oracle import oracle as O


def test_kvaerno5_order_conditions():
    A, b, bh, c = O.kvaerno5_tableau()
    assert np.allclose(A.sum(1), c, atol=1e-15)
    assert np.allclose(b, A[6])              # stiffly accurate
    assert np.all(np.diag(A)[1:] == 0.26)

    def conds(w, order):
        out = [(w.sum(), 1.0), (w @ c, 1 / 2), (w @ c ** 2, 1 / 3), (w @ A @ c, 1 / 6)]
        if order >= 4:
            out += [(w @ c ** 3, 1 / 4), (w @ (c * (A @ c)), 1 / 8), (w @ A @ c ** 2, 1 / 12), (w @ A @ A @ c, 1 / 24)]
        if order >= 5:
            out += [(w @ c ** 4, 1 / 5), (w @ (c ** 2 * (A @ c)), 1 / 10), (w @ ((A @ c) ** 2), 1 / 20),
                    (w @ (c * (A @ c ** 2)), 1 / 15), (w @ (c * (A @ A @ c)), 1 / 30), (w @ A @ c ** 3, 1 / 20),
                    (w @ A @ (c * (A @ c)), 1 / 40), (w @ A @ A @ c ** 2, 1 / 60), (w @ A @ A @ A @ c, 1 / 120)]
        return max(abs(x - y) for x, y in out)

    assert conds(b, 5) < 1e-15
    assert conds(bh, 4) < 1e-15
    assert conds(bh, 5) > 1e-4               # the embedded pair really is of lower order


@pytest.mark.parametrize("seed", [0, 1])
def test_spline_not_a_knot_matches_scipy(seed):
    from abcmb_amd import splines
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(-5, 0, 40))
    y = np.sin(3 * x) + rng.normal(0, 0.1, x.size)
    xq = np.linspace(-5.5, 0.5, 301)          # includes extrapolation on both sides
    ref = CubicSpline(x, y, bc_type="not-a-knot")(xq)
    assert np.max(np.abs(O.spline(x, y, xq) - ref)) < 1e-10
    assert np.max(np.abs(splines.evaluate(x, y, xq) - ref)) < 1e-10
    op = splines.operator(x, xq)
    assert np.max(np.abs(op @ y - ref)) < 1e-10


def _wigner_closed(l, m, n, mu):
    """d^l_mn(beta) with the reference's (-sin(beta/2))^(m-n) convention (m >= |n|)."""
    beta = np.arccos(mu)
    lognorm = 0.5 * (gammaln(l + m + 1) + gammaln(l - m + 1) - gammaln(l + n + 1) - gammaln(l - n + 1))
    return (np.exp(lognorm) * (-np.sin(beta / 2)) ** (m - n) * np.cos(beta / 2) ** (m + n)
            * eval_jacobi(l - m, m - n, m + n, mu))


@pytest.mark.parametrize("m,n", [(0, 0), (1, 1), (1, -1), (2, 2), (2, -2), (2, 0), (3, 1), (3, -3), (3, -1),
                                 (4, 0), (4, -4), (4, -2)])
def test_wigner_d_closed_form(m, n):
    mu = np.concatenate([np.linspace(-0.999, 0.999, 37), [1.0]])
    n_l = 80
    got = O.wigner(m, n, mu, n_l)
    for j in range(n_l):
        l = j + 2
        ref = _wigner_closed(l, m, n, mu) if l >= m else np.zeros_like(mu)
        assert np.max(np.abs(got[:, j] - ref)) < 1e-11, (l, m, n)


def test_bessel_tables_match_reference_data():
    """Regenerated phi0/phi1/phi2 tables equal the reference's shipped tables."""
    import os
    from abcmb_amd import bessel
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "bessel_ref_sample.npz"))
    ls, tab, rng = bessel.load_tables()
    assert np.array_equal(ls, g["l"])
    for f in range(3):
        got = tab[f][:, g["rows"]].T          # rows x l
        ref = g[f"phi{f}"]
        # the reference's text files carry 11 significant digits and a rounded x-grid
        assert np.max(np.abs(got - ref)) < 5e-10, f


def test_los_lookup_and_k_quadrature_vs_scipy():
    """A source at one conformal time turns the LOS into j_l(k chi); check against scipy."""
    from abcmb_amd import bessel
    ls, tab, rng = bessel.load_tables()
    kt = np.geomspace(1e-4, 0.3, 400)
    n_tau = 3
    tau_los = np.array([280.0, 300.0, 320.0])
    tau0 = 14000.0
    src = np.zeros((4, n_tau, kt.size))
    src[0, 1] = 1.0          # T0 at chi = 13700
    src[3, 1] = 1.0          # E at the same time
    ell_idx = np.arange(0, 60, 7, dtype=np.int32)
    cl, delta = O.los_cl(src, tau_los, tau0, kt, tab, rng, ls, ell_idx, 2e-9, 0.96, 0.05, want_delta=True)
    chi = tau0 - tau_los[1]
    for e, li in enumerate(ell_idx):
        l = int(ls[li])
        x = kt * chi
        j = spherical_jn(l, x)
        inside = (x >= rng[0, li, 0]) & (x < rng[0, li, 1])
        above = x >= rng[0, li, 1]
        assert np.max(np.abs(delta[0, e, inside] - j[inside])) < 2e-6
        # large-x branch: one-term Debye asymptotics, as in the reference
        if above.any():
            assert np.max(np.abs(delta[0, e, above] - j[above])) < 3e-3 * np.max(np.abs(j))
        eps = np.sqrt(3 / 8 * (l + 2) * (l + 1) * l * (l - 1))
        E = np.where(x > 0, delta[0, e] / np.where(x > 0, x, 1) ** 2 * eps, 0.0)
        assert np.allclose(delta[3, e], E, rtol=1e-12, atol=1e-300)
        P = 4 * np.pi * 2e-9 * (kt / 0.05) ** (0.96 - 1) / kt
        T = delta[0, e] + delta[1, e] + delta[2, e]
        assert math.isclose(cl[0, e], np.trapezoid(P * T * T, kt), rel_tol=1e-12)
        assert math.isclose(cl[2, e], np.trapezoid(P * delta[3, e] ** 2, kt), rel_tol=1e-12)


def test_zero_lensing_potential_is_identity():
    n_l = 250
    ls = np.arange(2, n_l + 2)
    rng = np.random.default_rng(3)
    tt = 1e-10 / (ls * (ls + 1.0)) * (1 + 0.3 * rng.random(n_l))
    ee = 1e-12 / (ls * (ls + 1.0)) * (1 + 0.3 * rng.random(n_l))
    te = 0.5 * np.sqrt(tt * ee) * np.cos(ls / 7.0)
    mu, w = np.polynomial.legendre.leggauss(n_l + 2 + 70)
    mu = np.concatenate([mu, [1.0]]); w = np.concatenate([w, [0.0]])
    out = O.lens(np.stack([tt, te, ee]), np.zeros(n_l), mu, w)
    # exact up to the cancellation of the l-sums in xi(mu)
    assert np.max(np.abs(out[0] / tt - 1)) < 1e-8
    assert np.max(np.abs(out[2] / ee - 1)) < 1e-8
    assert np.max(np.abs(out[1] - te)) < 1e-9 * np.max(np.abs(te))
    assert np.max(np.abs(out[3])) < 1e-9 * np.max(ee)


def test_lensing_generates_positive_bb():
    n_l = 400
    ls = np.arange(2, n_l + 2).astype(float)
    tt = 1e-9 / (ls * (ls + 1.0))
    ee = 1e-11 * np.ones(n_l) / (ls + 10)
    te = 0.3 * np.sqrt(tt * ee)
    clpp = 1e-7 * (ls * (ls + 1.0)) ** -2 * 2 * np.pi
    mu, w = np.polynomial.legendre.leggauss(n_l + 2 + 70)
    mu = np.concatenate([mu, [1.0]]); w = np.concatenate([w, [0.0]])
    out = O.lens(np.stack([tt, te, ee]), clpp, mu, w)
    assert np.all(out[3][10:300] > 0)
    assert np.max(np.abs(out[0] / tt - 1)[:300]) < 0.2


def test_branch_frozen_los():
    """``oracle.los_cl_branch``: with the branch distances equal to the actual ones
    it is ``los_cl`` bit for bit; with them frozen at a base tau0, central
    differences in tau0 converge: against a step of 1e-5 Mpc, steps of 1e-2 and
    1e-4 Mpc differ by 2.4e-5 and 2.4e-6 of the derivative's maximum per
    spectrum (first order in the step: the table is interpolated linearly, so
    the differences see the slope changes between segments), the property the
    tangent-LOS check on the GPU relies on."""
    from abcmb_amd import bessel
    ls, tab, rng = bessel.load_tables()
    kt = np.geomspace(1e-4, 0.5, 600)
    tau_los = np.linspace(250.0, 13500.0, 60)
    tau0 = 14000.0
    g = np.random.default_rng(3)
    shape_t = np.exp(-((tau_los - 300.0) / 60.0) ** 2) + 0.05 * np.exp(-((tau_los - 9000.0) / 3000.0) ** 2)
    src = np.stack([g.uniform(0.5, 1.5) * shape_t[:, None] * np.exp(-(kt[None, :] / 0.2) ** 2)
                    * (1.0 + 0.1 * q * np.log(kt[None, :])) for q in range(4)])
    ell_idx = np.arange(0, ls.size, 8, dtype=np.int32)
    args = (kt, tab, rng, ls, ell_idx, 2e-9, 0.96, 0.05)
    chi = tau0 - tau_los
    assert np.array_equal(O.los_cl_branch(src, tau_los, tau0, *args, chi), O.los_cl(src, tau_los, tau0, *args))

    def d(h):
        f = lambda t0: O.los_cl_branch(src, tau_los, t0, *args, chi)
        return (f(tau0 + h) - f(tau0 - h)) / (2 * h)

    ref = d(1e-5)
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.max(np.abs(d(1e-2) - ref) / scale) < 5e-5
    assert np.max(np.abs(d(1e-4) - ref) / scale) < 5e-6
