"""Forward-mode derivatives of the host producer (``Model.tangent_plans``,
``abcmb_host_background_tangent`` and its LOS / evaluation companions).

Pins, on a cosmology with a massive neutrino and lensing:
  - the value parts of the dual evaluation equal the producer's tables bit for
    bit (``Background.tangents`` raises otherwise);
  - the forward derivatives are the limit of the frozen-grid central
    differences: for x_e, tau(ln a), the LOS background vectors and the
    lensing background, the difference between central differences at a
    relative step h and the forward derivative falls as h^2 (ratio ~100 from
    h = 1e-3 to 1e-4) down to the differences' rounding floor, and is below
    1e-6 of each derivative's maximum at h = 1e-5;
  - the reionization redshift's derivative (implicit-function theorem on the
    tau_reion root) against the central difference of the root;
  - closed forms of the input tangents: d omega_b / d omega_b = 1,
    d omega_Lambda / d h = 2 h, zeros where a parameter does not enter;
  - a model with user fluids (w0wa dark energy in place of the cosmological
    constant): the forward derivatives through the generated host code against
    fine central differences (tests/test_fluids.py).
"""
import numpy as np
import pytest

from abcmb_amd import abi
from abcmb_amd import likelihood as LK

FIELDS = [f for f, _ in abi.AbcmbCosmo._fields_]


@pytest.fixture(scope="module")
def setup():
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=500, k_axis_perturbations=np.geomspace(1e-4, 0.3, 4))
    th = LK.FIDUCIAL.copy()
    th[6] = 0.12
    base = m.prepare(m.derived(LK.theta_to_params(th)))

    def plans(rel, forward):
        st = rel * np.abs(th)
        q = LK._pm_params(m, th, st)
        sd = [st[p] for p in LK.ODE_DIRS]
        pm = m.tangent_plans(base, q, sd) if forward else [m.prepare(x, freeze=base) for x in q]
        return pm, sd

    fwd, sd = plans(1e-4, True)
    fd = {rel: plans(rel, False) for rel in (1e-3, 1e-4, 1e-5)}
    seeds = lambda pm, s: tangent_seeds(pm[0::2], pm[1::2], s)
    return m, th, base, fwd, sd, fd, seeds


def _diff(pm, sd, get):
    return np.stack([(get(pm[2 * d]) - get(pm[2 * d + 1])) / (2 * sd[d]) for d in range(len(sd))])


def _err(a, b):
    """max |a - b| per direction, relative to max |b| of that direction."""
    a, b = a.reshape(a.shape[0], -1), b.reshape(b.shape[0], -1)
    sc = np.abs(b).max(axis=1)
    return np.where(sc > 0, np.abs(a - b).max(axis=1) / np.where(sc > 0, sc, 1.0), 0.0), sc


@pytest.mark.parametrize("what", ["xe", "tau", "bgv", "lens_aH", "lens_tau"])
def test_forward_is_the_limit_of_central_differences(setup, what):
    m, th, base, fwd, sd, fd, seeds = setup
    get = {"xe": lambda p: p.bg.xe_tab, "tau": lambda p: p.bg.tau_tab, "bgv": lambda p: p.bgv,
           "lens_aH": lambda p: p.lens_aH, "lens_tau": lambda p: p.lens_tau}[what]
    f = _diff(fwd, sd, get)
    e3, sc = _err(_diff(*fd[1e-3], get), f)
    e4, _ = _err(_diff(*fd[1e-4], get), f)
    e5, _ = _err(_diff(*fd[1e-5], get), f)
    for d in range(len(sd)):
        if sc[d] == 0.0:
            continue
        assert e5[d] < 1e-6, (what, d, e3[d], e4[d], e5[d])
        if e4[d] > 1e-8:
            # truncation dominates: second order
            assert 30.0 < e3[d] / e4[d] < 300.0, (what, d, e3[d], e4[d])


def test_tm_and_fields_against_fine_differences(setup):
    m, th, base, fwd, sd, fd, seeds = setup
    sf, s5 = seeds(fwd, sd), seeds(*fd[1e-5])
    e, sc = _err(sf.tm, s5.tm)
    big = sc > 1e-6 * sc.max()
    assert np.all(e[big] < 1e-5), e
    # the double fields the producer computes: tau0, x_e and T_m today
    for name in ("tau0", "xe_last", "tm_last"):
        i = FIELDS.index(name)
        a, b = sf.par[:, i], s5.par[:, i]
        assert np.all(np.abs(a - b) <= 1e-6 * np.abs(b).max() + 1e-14), (name, a, b)


def test_zreion_derivative_by_implicit_function(setup):
    m, th, base, fwd, sd, fd, seeds = setup
    i = FIELDS.index("z_reion")
    d = LK.ODE_DIRS.index(LK.NAMES.index("tau_reion"))
    f = (fwd[2 * d].bg.cosmo.z_reion - fwd[2 * d + 1].bg.cosmo.z_reion) / (2 * sd[d])
    pm, s = fd[1e-4]
    c = (pm[2 * d].bg.cosmo.z_reion - pm[2 * d + 1].bg.cosmo.z_reion) / (2 * s[d])
    assert f > 0 and abs(f / c - 1) < 1e-6, (f, c)
    # omega_b moves the root too (through n_H)
    db = (fwd[0].bg.cosmo.z_reion - fwd[1].bg.cosmo.z_reion) / (2 * sd[0])
    cb = (pm[0].bg.cosmo.z_reion - pm[1].bg.cosmo.z_reion) / (2 * s[0])
    assert abs(db / cb - 1) < 1e-5, (db, cb)


def test_input_tangent_closed_forms(setup):
    m, th, base, fwd, sd, fd, seeds = setup
    par = seeds(fwd, sd).par
    ob, oc, h, tr, mnu = range(5)
    col = FIELDS.index
    assert abs(par[ob, col("omega_b")] - 1.0) < 1e-10
    assert abs(par[oc, col("omega_cdm")] - 1.0) < 1e-10
    assert abs(par[h, col("h")] - 1.0) < 1e-10
    hh = th[2]
    assert abs(par[h, col("omega_Lambda")] / (2 * hh) - 1.0) < 1e-8
    assert abs(par[mnu, col("m_ncdm")] - 1.0) < 1e-10
    for d in (ob, oc, mnu):
        assert par[d, col("h")] == 0.0
    assert np.all(par[tr, :col("tau0")] == 0.0)
    # grid metadata is held
    for name in ("xe_x0", "xe_dx", "xe_lna_last", "tm_x0", "lna_rec", "lna_transfer_start"):
        assert np.all(par[:, col(name)] == 0.0), name


def test_plans_keep_the_base_grids(setup):
    m, th, base, fwd, sd, fd, seeds = setup
    for p in fwd:
        assert p.lna_save is base.lna_save and p.lna_start is base.lna_start
        assert np.array_equal(p.bg.rec_switch, base.bg.rec_switch)
        assert p.bg.cosmo.lna_rec == base.bg.cosmo.lna_rec
        assert p.bg.cosmo.tau_tab == p.bg.tau_tab.ctypes.data


def test_host_tangents_switch():
    import types
    assert not LK._forward_host(types.SimpleNamespace(host_tangents="fd"))
    assert LK._forward_host(types.SimpleNamespace(host_tangents="forward"))
