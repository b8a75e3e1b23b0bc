"""CPU tests of the perturbation-ODE oracle, the host background producer,
the grids and the derived parameters."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from abcmb_amd import abi, constants as cnst
from abcmb_amd.model import Model
from abcmb_amd.params import species_rho
from oracle import oracle as O
from oracle import pipeline


@pytest.fixture(scope="module")
def c1(fiducial):
    m = Model(l_max=1000, l_max_g=12, l_max_pol_g=12, l_max_massless_nu=12,
              k_axis_perturbations=np.geomspace(1e-4, 0.3, 200))
    p = m.derived(fiducial)
    return m, p, m.prepare(p)


def test_reference_grids():
    # PT k-grid sizes recorded by the reference (ABCMB_basics.ipynb:516) and the
    # transfer grid / sampled multipoles of abcmb/model_specs.py, abcmb/spectrum.py
    m = Model()
    assert m.k_pt.size == 548 and m.n_save == 500
    assert m.k_transfer.size == 1704 and m.ell_idx.size == 99
    ml = Model(lensing=True)
    assert ml.k_pt.size == 571 and ml.ell_idx.size == 111 and ml.mu.size == 3071
    m1 = Model(l_max=1000)
    assert m1.k_transfer.size == 763 and m1.ell_idx.size == 61


def test_derived_parameters(fiducial, c1):
    m, p, _ = c1
    # N_eff = 3.044 with T_nu = 0.71636856 T_cmb is three neutrinos (abcmb/main.py:573-587)
    assert abs(p["N_nu_massless"] - 3.0) < 1e-5
    assert abs(p["omega_Lambda"] - (p["h"] ** 2 - p["omega_r"] - p["omega_m"])) < 1e-15
    assert abs(p["R_nu"] - 0.4087) < 1e-3
    with pytest.raises(ValueError):
        m.derived(dict(fiducial, N_nu_massless=3))


def test_host_background(c1):
    m, p, plan = c1
    bg = plan.bg
    # conformal time against a direct quadrature of 1/aH
    from scipy.integrate import quad
    ev = lambda l: bg.eval(np.array([l]))["aH"][0]
    t_cut = math.exp(-16.1) / (cnst.H0_over_h / cnst.c_Mpc_over_s) / math.sqrt(p["omega_r"])
    for lna in (-10.0, -6.0, -1.0):
        ref = t_cut + quad(lambda l: 1.0 / ev(l), -16.1, lna, epsabs=0, epsrel=1e-12, limit=200)[0]
        assert abs(bg.eval(np.array([lna]))["tau"][0] / ref - 1) < 1e-6
    assert 14000 < bg.tau0 < 14400
    assert 1080 < math.exp(-bg.lna_rec) - 1 < 1100
    assert -8 < bg.lna_transfer_start < -6.5
    assert 6 < bg.z_reion < 9
    e = bg.eval(np.array([-12.0, -math.log(1101), -math.log(21)]))
    assert abs(e["xe"][0] - (1 + 2 * p["YHe"] * cnst.mH / cnst.mHe / (1 - p["YHe"]))) < 1e-3
    assert 0.05 < e["xe"][1] < 0.3 and e["xe"][2] < 1e-3


def test_einstein_equations_hold(c1):
    """MB95 (21c) and (21d) are not used by the evolution; the RHS must satisfy them."""
    m, p, plan = c1
    c, st = plan.bg.cosmo, m.settings
    lay = m.layout
    rng = np.random.default_rng(7)
    G, cM = cnst.G, cnst.c_Mpc_over_s
    for k, lna in [(0.01, -7.0), (0.2, -3.0), (0.05, -0.5)]:
        y = rng.normal(size=lay.n)
        dy = O.rhs(c, st, k, lna, y)
        eps = 1e-5
        dt = (O.rhs(c, st, k, lna + eps, y) - O.rhs(c, st, k, lna - eps, y)) / (2 * eps)
        dtot = dt + O.rhs(c, st, k, lna, dy)
        H, Hl = -2 * dy[1], -2 * dtot[1]
        ep, epp = dy[0], dtot[0]
        aH = plan.bg.eval(np.array([lna]))["aH"][0]
        aHp = (plan.bg.eval(np.array([lna + eps]))["aH"][0] - plan.bg.eval(np.array([lna - eps]))["aH"][0]) / (2 * eps)
        hd, hdd = aH * H, aH * (aHp * H + aH * Hl)
        ed, edd = aH * ep, aH * (aHp * ep + aH * epp)
        a = math.exp(lna)
        K = 4 * math.pi * G * a * a / cM ** 2
        rg, ru = species_rho("Photon", lna, p), species_rho("MasslessNeutrino", lna, p)
        dP = (rg * y[lay.i_g] + ru * y[lay.i_ur]) / 3.0
        sig = 4.0 / 3.0 * (rg * y[lay.i_g + 2] + ru * y[lay.i_ur + 2])
        eta = y[0]
        t21c = [hdd, 2 * aH * hd, -2 * k * k * eta, 6 * K * dP]
        t21d = [hdd, 6 * edd, 2 * aH * (hd + 6 * ed), -2 * k * k * eta, 6 * K * sig]
        assert abs(sum(t21c)) < 1e-6 * max(abs(v) for v in t21c), (k, lna, t21c)
        assert abs(sum(t21d)) < 1e-6 * max(abs(v) for v in t21d), (k, lna, t21d)


def test_oracle_integrator_converges_to_radau(c1):
    m, p, plan = c1
    c = plan.bg.cosmo
    st = abi.AbcmbOdeSettings.from_buffer_copy(m.settings)
    st.max_steps = 50000
    lay = m.layout
    lna_save = np.ascontiguousarray(plan.lna_save[::50])
    kk = np.array([2e-4, 0.006, 0.03])
    ls = plan.bg.lna_start(kk, 0.0015, 0.07)
    cc = abi.AbcmbCosmo.from_buffer_copy(c)
    pt = np.zeros((abi.N_PT_FIELDS, lna_save.size, kk.size))
    cc.lna_save, cc.k, cc.pt, cc.n_save, cc.nk = lna_save.ctypes.data, kk.ctypes.data, pt.ctypes.data, lna_save.size, kk.size
    cos = (abi.AbcmbCosmo * 1)(cc)
    items = (abi.AbcmbPertItem * 3)(*[abi.AbcmbPertItem(k=kk[i], lna_start=ls[i], rtol=1e-9, atol=1e-13, cosmo=0, ik=i)
                                      for i in range(3)])
    status, steps, states = O.pert(cos, items, st, want_states=True)
    assert np.all(status == 0)
    key = [0, 1, 2, 3, lay.i_g, lay.i_g + 1, lay.i_g + 2, lay.i_pol, lay.i_ur, lay.i_ur + 2]
    for i in range(3):
        sol = solve_ivp(lambda t, y: O.rhs(c, st, kk[i], t, y), (ls[i], 0.0), O.ic(c, st, kk[i], ls[i]),
                        method="Radau", rtol=1e-11, atol=1e-15, t_eval=lna_save)
        ref = sol.y.T
        for j in key:
            scale = np.abs(ref[:, j]).max()
            assert np.max(np.abs(states[i][:, j] - ref[:, j])) < 2e-6 * scale, (kk[i], j)


def test_oracle_reports_max_steps(c1):
    m, p, plan = c1
    st = abi.AbcmbOdeSettings.from_buffer_copy(m.settings)
    st.max_steps = 20
    saved = m.settings
    m.settings = st
    try:
        out = pipeline.run(m, plan, k_subset=[150])
    finally:
        m.settings = saved
    assert out["status"][0] == 1 and out["steps"][0] == 20


def test_oracle_pipeline_config1_physics(c1):
    """Config 1 end to end on the oracle: spectrum shape checks."""
    m, p, plan = c1
    out = pipeline.run(m, plan)
    assert np.all(out["status"] == 0)
    cl = out["cl_full"]
    l = m.lensing_ells.astype(float)
    dtt = l * (l + 1) * cl[0] / (2 * np.pi) * (2.7255e6) ** 2
    assert np.all(cl[0] > 0) and np.all(cl[2] > 0)
    assert 700 < dtt[10 - 2] < 1300               # Sachs-Wolfe plateau, uK^2
    ipk = np.argmax(dtt[100:400]) + 100
    assert 200 < l[ipk] < 240 and 4500 < dtt[ipk] < 7000   # first acoustic peak
    assert cl[1][150 - 2] < 0                      # TE anticorrelation near l = 150
    assert np.all(out["pk"] > 0)
