"""The host producer's recombination (HyRex restated in csrc/hyrex.cpp) against
the oracle's independent restatement (oracle/hyrex.py: scipy Radau at
rtol 1e-11, scipy's map_coordinates, central-difference post-Saha
derivatives), on the oracle's own background (H from oracle_bg, n_H and T_CMB
from the parameters): the same stage hand-overs, x_e and T_m within 1e-8 of
the oracle's (the product integrates on the save grid with an order-3 SDIRK: measured
<= 5e-9);
plus the limits the physics fixes."""
import math

import numpy as np
import pytest

from abcmb_amd import abi, constants as K, likelihood as LK
from abcmb_amd.background import Background
from abcmb_amd.model import Model, MassiveNeutrino
from oracle import hyrex as HX
from oracle import oracle as O


def _inputs(c):
    lna = HX.X0 + HX.DX * np.arange(HX.N_GRID)
    a = np.exp(lna)
    T = c.TCMB0 / a
    nh = (1 - c.YHe) * c.omega_b * K.RHO_UNIT / K.mH / a ** 3
    H = np.array([O.bg(c, l)["aH"] for l in lna]) * K.c_Mpc_over_s / a
    return T, nh, H


@pytest.mark.parametrize("massive,seed", [(False, 0), (True, 1)])
def test_recombination_matches_independent_restatement(massive, seed):
    m = Model(user_species=(MassiveNeutrino,) if massive else (), l_max=1000)
    th = LK.latin_hypercube(1, seed=seed)[0]
    p = m.derived(LK.theta_to_params(th))
    b = Background(p, massive, reionization=False)
    c = b.cosmo
    xe_o, tm_o, jt0, ends = HX.history(*_inputs(c), c.YHe, c.omega_b, c.omega_cdm, p["Neff"])
    assert list(b.rec_switch) == ends
    n = int(round((c.xe_lna_last - c.xe_x0) / c.xe_dx)) + 1
    assert n == xe_o.size
    assert c.tm_x0 == pytest.approx(HX.X0 + jt0 * HX.DX, abs=1e-12)
    np.testing.assert_allclose(b.xe_tab[:n], xe_o, rtol=1e-8, atol=0)
    assert np.all(b.xe_tab[n:] == b.xe_tab[n - 1])
    m_t = int(round((c.tm_lna_last - c.tm_x0) / c.tm_dx)) + 1
    np.testing.assert_allclose(b.tm_tab[:m_t], tm_o, rtol=1e-8, atol=0)
    assert b.xe_tab.size == abi.C.c_int(28925).value and b.tm_tab.size == 16384


def test_recombination_limits():
    """Fully ionized (H+ and He++) at z = 8000; x_e(z ~ 1100) ~ 0.1-0.2 and freeze-out
    at ~2e-4 (the textbook history); the matter temperature tracks the radiation's
    where the table starts (z ~ 700) and falls below it by z = 0."""
    m = Model(l_max=1000)
    p = m.derived(LK.theta_to_params(LK.FIDUCIAL))
    b = Background(p, False, reionization=False)
    c = b.cosmo
    fhe = c.YHe / (1 - c.YHe) / 3.97153
    assert b.xe_tab[0] == pytest.approx(1 + 2 * fhe, rel=1e-3)
    j = int(round((-math.log(1101.0) - c.xe_x0) / c.xe_dx))
    assert 0.1 < b.xe_tab[j] < 0.2
    assert 1e-4 < c.xe_last < 3e-4
    T0 = c.TCMB0 / math.exp(c.tm_x0)
    assert b.tm_tab[0] == pytest.approx(T0, rel=1e-2)   # Compton coupling: a 0.3% lag at z ~ 700
    assert b.tm_tab[0] < T0
    assert c.tm_last < 0.1 * c.TCMB0
