"""The oracle's tangent-linear integration, pinned against derivatives it does
not compute itself.

The tangent oracle (``pipeline.tangent_tables``) integrates the perturbation
table's derivatives along the parameter directions at the plain
integration's step sequence, each stage tangent solved to rounding.  Pins:
  - its values are the plain oracle's, bit for bit (the tangent rides along
    without touching the value path);
  - a direction that does not reach the equations (m_nu without a massive
    species) gives exactly zero;
  - central differences of *replays* of the recorded step sequence at
    theta +- h with every stage solved to rounding ("converged" mode: the
    derivative of the implicit Runge-Kutta map at frozen steps, computed by
    finite differences of whole integrations) agree to 5e-3 of each field's
    maximum, 3e-4 in the median.  The two differ by the derivative of the
    value stages' convergence error (the chord iteration stops at kappa =
    1e-2 of the error tolerance; the tangents are taken at those stage
    values, as an implicit-function derivative of the stage equations is,
    the replay at the exactly solved ones): a discretization effect of the
    size of the tolerance, largest where a field's derivative is small
    against the field itself.  With 100x tighter tolerances the median
    drops by more than ten (to below 2e-5);
  - the replay machinery itself: replaying at the base point reproduces the
    recorded integration bit for bit, and the replayed differences do not
    change with the difference step (1e-4 .. 4e-4 relative) beyond 1e-5.
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK


def _setup(massive, nk=6, tol=1.0):
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.3, nk)
    kw = {}
    if tol != 1.0:
        kw = dict(rtol_small_k_PE=1e-5 * tol, rtol_large_k_PE=1e-4 * tol, atol_small_k_PE=1e-10 * tol,
                  atol_large_k_PE=1e-6 * tol, max_steps_PE=20000)
    m = Model(user_species=(MassiveNeutrino,) if massive else (), l_max=500, k_axis_perturbations=kk, **kw)
    th = LK.FIDUCIAL.copy()
    base = m.prepare(m.derived(LK.theta_to_params(th)))
    return m, th, base


def _pm(m, th, base, rel=LK.FD_REL_STEP):
    steps = rel * np.abs(th)
    plus, minus = [], []
    for p in LK.ODE_DIRS:
        for sgn, lst in ((1, plus), (-1, minus)):
            t = th.copy()
            t[p] += sgn * steps[p]
            lst.append(m.prepare(m.derived(LK.theta_to_params(t)), freeze=base))
    return plus, minus, [steps[p] for p in LK.ODE_DIRS]


def _rel(a, b):
    """max |a - b| per (direction, field), relative to each field's max |b|."""
    sc = np.abs(b).max(axis=(-2, -1)) + 1e-300
    return np.abs(a - b).max(axis=(-2, -1)) / sc


@pytest.mark.parametrize("massive,tol,median", [(False, 1.0, 3e-4), (True, 1.0, 3e-4), (False, 1e-2, 2e-5)])
def test_tangent_oracle_against_converged_replays(massive, tol, median):
    from oracle import pipeline
    m, th, base = _setup(massive, tol=tol)
    plus, minus, sd = _pm(m, th, base)
    ks = np.arange(m.k_pt.size)
    pt, dpt, nst = pipeline.tangent_tables(m, base, plus, minus, sd, ks)
    plain = pipeline.run(m, base, k_subset=ks)
    assert np.array_equal(pt, plain["pt"]) and np.array_equal(nst, plain["steps"])
    _, ref = pipeline.frozen_step_tangents(m, base, plus, minus, sd, ks, mode="converged")
    e = _rel(dpt, ref)
    active = [d for d, p in enumerate(LK.ODE_DIRS) if massive or LK.NAMES[p] != "m_nu"]
    assert np.max(e[active]) < 5e-3, e
    assert np.median(e[active]) < median, e
    if not massive:
        assert np.all(dpt[LK.ODE_DIRS.index(6)] == 0.0)   # m_nu does not reach a massless model


def test_replay_is_smooth_and_reproduces_the_record():
    from oracle import pipeline
    m, th, base = _setup(False, nk=4)
    ks = np.arange(4)
    pt0, d1 = pipeline.frozen_step_tangents(m, base, *_pm(m, th, base, 1e-4), ks, mode="converged")
    assert np.array_equal(pt0, pipeline.run(m, base, k_subset=ks)["pt"])
    _, d2 = pipeline.frozen_step_tangents(m, base, *_pm(m, th, base, 2e-4), ks, mode="converged")
    _, d3 = pipeline.frozen_step_tangents(m, base, *_pm(m, th, base, 4e-4), ks, mode="converged")
    # a smooth map: the differences do not depend on the step (no step-sequence jumps)
    assert np.median(_rel(d2[:4], d1[:4])) < 1e-5
    assert np.median(_rel(d3[:4], d2[:4])) < 1e-5
