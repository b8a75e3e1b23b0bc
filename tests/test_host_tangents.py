"""Tangent seeds of the host producer (``model.tangent_seeds``).

The seeds are central differences of the producer between plans at
theta +- h with every grid frozen.  Pins:
  - closed forms of the derived parameters: d omega_b / d omega_b = 1,
    d omega_Lambda / d omega_b = -1, d R_b / d omega_b = (omega_m - omega_b) /
    omega_m^2, d om / d omega_b = H0 / (c sqrt(omega_r)), d omega_Lambda / d h =
    2 h, d m_ncdm / d m_nu = 1, and zeros where a parameter does not enter;
  - smoothness of the producer with frozen grids: the seeds of every field and
    table do not change with the difference step (1e-4 vs 2e-4 relative) beyond
    what central differences allow, also for tau_reio, which goes through the
    reionization root (z_reion) and the x_e table (measured: 1e-6 .. 1e-10,
    shrinking as h^2).
"""
import math

import numpy as np

from abcmb_amd import abi, constants as cnst
from abcmb_amd import likelihood as LK

FIELDS = [f for f, _ in abi.AbcmbCosmo._fields_[:abi.COSMO_DOUBLES]]


def _seeds(rel):
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    m = Model(user_species=(MassiveNeutrino,), l_max=500, k_axis_perturbations=np.geomspace(1e-4, 0.3, 4))
    th = LK.FIDUCIAL.copy()
    base = m.prepare(m.derived(LK.theta_to_params(th)))
    steps = rel * np.abs(th)
    plus, minus = [], []
    for p in LK.ODE_DIRS:
        for sgn, lst in ((1, plus), (-1, minus)):
            t = th.copy()
            t[p] += sgn * steps[p]
            lst.append(m.prepare(m.derived(LK.theta_to_params(t)), freeze=base))
    return base, tangent_seeds(plus, minus, [steps[p] for p in LK.ODE_DIRS])


def test_seed_closed_forms():
    base, sd = _seeds(1e-4)
    P = base.params
    d_b, d_c, d_h, d_tau, d_m = range(5)
    col = {f: i for i, f in enumerate(FIELDS)}
    par = sd.par
    assert abs(par[d_b, col["omega_b"]] - 1.0) < 1e-9
    assert abs(par[d_b, col["omega_Lambda"]] + 1.0) < 1e-7
    wm = P["omega_m"]
    assert abs(par[d_b, col["R_b"]] - (wm - P["omega_b"]) / wm ** 2) < 1e-7 * abs((wm - P["omega_b"]) / wm ** 2)
    # omega_r sums every species at a = e^-23, baryons included (abcmb/main.py), which moves
    # d om / d omega_b by (omega_m / omega_r) e^-23 / 2 ~ 2e-7 relative from the closed form
    d_om = cnst.H0_over_h / cnst.c_Mpc_over_s / math.sqrt(P["omega_r"])
    assert abs(par[d_b, col["om"]] - d_om) < 1e-6 * d_om
    assert abs(par[d_h, col["omega_Lambda"]] - 2 * P["h"]) < 1e-7
    assert abs(par[d_h, col["h"]] - 1.0) < 1e-9
    assert abs(par[d_m, col["m_ncdm"]] - 1.0) < 1e-9
    # parameters that do not enter stay fixed (N_ur only through the matter density the
    # reference's N_eff bookkeeping counts at early times, ~1e-10)
    for f in ("A_s", "n_s", "k_pivot", "TCMB0", "YHe", "T_ur", "T_ncdm", "xe_x0", "xe_dx", "lna_rec",
              "lna_transfer_start"):
        assert np.all(par[:, col[f]] == 0.0), f
    assert np.all(np.abs(par[:, col["N_ur"]]) < 1e-8)
    assert np.all(par[[d_b, d_c, d_h, d_m], col["z_reion"]] != 0.0) and par[d_tau, col["z_reion"]] > 0.0
    assert np.all(sd.tau[d_tau] == 0.0)            # tau_reio does not move the conformal-time table


def test_seeds_are_smooth_in_the_step():
    """Central differences at 1e-4 and 2e-4 agree to 1e-5 of each seed's maximum
    (T_m: 1e-3 -- its derivative along h is ~1e-9 of T_m itself, so the
    recombination integration's rounding, ~1e-15 absolute, is what is left)."""
    _, s1 = _seeds(1e-4)
    _, s2 = _seeds(2e-4)
    for a, b, bound in ((s1.par, s2.par, 1e-5), (s1.tau, s2.tau, 1e-5), (s1.xe, s2.xe, 1e-5),
                        (s1.tm, s2.tm, 1e-3)):
        scale = np.abs(a).max(axis=1, keepdims=True) + 1e-300
        assert np.max(np.abs(a - b) / scale) < bound
