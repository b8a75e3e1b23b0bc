"""User-defined fluids on the CPU: the traced coefficients against the fluids'
own methods called with numbers, the oracle's fluid callbacks against the
standard species they restate, the w0wa fluid at w = -1 against LCDM, and the
forms the kernels refuse.

Models with user fluids build their host producer here (g++, seconds); no
GPU library is touched.
"""
import math

import numpy as np
import pytest

from abcmb_amd import abi, examples as EX, fluidgen as FG, species as SP
from abcmb_amd.model import Model
from oracle import oracle as O

SMALL = dict(l_max=1000, l_max_g=12, l_max_pol_g=12, l_max_massless_nu=12, max_steps_PE=4000,
             k_axis_perturbations=np.geomspace(2e-4, 0.2, 6))
P0 = dict(h=0.6736, omega_b=0.02237, omega_cdm=0.12, A_s=2.1e-9, n_s=0.9649, tau_reion=0.0544, Neff=3.044)


class CDMClone(SP.StandardFluid):
    """Cold dark matter again, with its own density parameter omega_x."""
    name = "CDMClone"
    num_equations = 1
    is_matter = True

    def rho(self, lna, params):
        return params["omega_x"] * SP.RHO_UNIT / SP.jnp.exp(lna) ** 3

    def P(self, lna, params):
        return 0.0

    def y_ini(self, k, tau, params):
        return SP.jnp.array([-(k * tau) ** 2 / 4.0 * (1.0 - params["om"] * tau / 5.0)])

    def y_prime(self, k, lna, hp, ep, y, args):
        return SP.jnp.array([-hp / 2.0])


def _env(model, p, lna, k, aH, tau):
    env = {"lna": lna, "a": math.exp(lna), "k": k, "aH": aH, "tau": tau}
    KG = 4.0 * math.pi * 1.18980205e-40 / (9.71561e-15 ** 2)
    a = math.exp(lna)
    env["kh"] = 2.0 * KG * a * a / (aH * aH)
    env["ke"] = KG * a * a / (aH * k * k)
    for name, fld in FG.PARAM_FIELDS.items():
        env[("field", fld)] = p[name]
    for j, name in enumerate(model.fluids.params):
        env[("fluid", j)] = p[name]
    return env


def _nadm():
    return Model(user_species=(SP.DarkEnergy, EX.InteractingDM, EX.InteractingDR, SP.Baryon, SP.Photon,
                               SP.MasslessNeutrino), use_LCDM_species=False, **SMALL)


def _w0wa():
    return Model(user_species=(SP.ColdDarkMatter, SP.Baryon, SP.Photon, SP.MasslessNeutrino, EX.W0WaDarkEnergy),
                 use_LCDM_species=False, **SMALL)


@pytest.mark.parametrize("which", ["sidr", "nadm", "w0wa"])
def test_traced_coefficients_match_the_methods(which):
    """Every row, weight, density and initial condition the generator extracted,
    evaluated at random points, equals the fluid's own methods called with
    numbers (the reference's calling convention, abcmb/perturbations.py:201-248)."""
    if which == "sidr":
        m, extra = Model(user_species=(EX.SelfInteractingDR,), **SMALL), dict(N_idr=0.3)
    elif which == "nadm":
        m, extra = _nadm(), dict(Gamma_0_nadm=1e-3, N_idr=0.2)
    else:
        m, extra = _w0wa(), dict(w0DE=-0.8, waDE=-0.6)
    fs = m.fluids
    p = m.derived(dict(P0, **extra))
    rng = np.random.default_rng(3)
    tail = (m.species_objs, {o.name: i for i, o in enumerate(m.species_objs)})
    for _ in range(5):
        lna, k, aH, tau = rng.uniform(-12, 0), 10 ** rng.uniform(-4, 0), 10 ** rng.uniform(-5, -1), rng.uniform(1, 300)
        env = _env(m, p, lna, k, aH, tau)
        rho = sum(float(f.rho(lna, p)) for f in fs.user)
        pres = sum(float(f.P(lna, p)) for f in fs.user)
        assert fs.rho.eval(env) == pytest.approx(rho, rel=1e-13)
        assert fs.pres.eval(env) == pytest.approx(pres, rel=1e-13, abs=1e-300)
        if fs.n_rows == 0:
            continue
        n = m.settings.n
        y = rng.standard_normal(n)
        hp, ep = rng.standard_normal(2)
        want = np.concatenate([np.asarray(f.y_prime(k, lna, hp, ep, y, (FG._NumBG(lna, aH), p) + tail))
                               for f in fs.user if f.num_equations])
        slots = [e.eval(env) for e in fs.slots]
        got = []
        for j in range(fs.n_rows):
            v = sum(c.eval(env) * (slots[sl] if sl >= 0 else 1.0) * y[fs.first_row + col]
                    for col, c, sl in fs.row_terms[j])
            v += fs.row_metric[j][0].eval(env) * hp + fs.row_metric[j][1].eval(env) * ep
            got.append(v)
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
        # weights of the user rows in h' and eta': kh rho_delta, ke (rho + P) theta
        for j in range(fs.n_rows):
            e = np.zeros(n)
            e[fs.first_row + j] = 1.0
            rd = sum(float(f.rho_delta(lna, e, p)) for f in fs.user if f.num_equations)
            rt = sum(float(f.rho_plus_P_theta(lna, e, p)) for f in fs.user if f.num_equations)
            hv = slots[fs.h_slot[j]] if fs.h_slot[j] >= 0 else 0.0
            ev = slots[fs.e_slot[j]] if fs.e_slot[j] >= 0 else 0.0
            assert hv == pytest.approx(env["kh"] * rd, rel=1e-13, abs=1e-300)
            assert ev == pytest.approx(env["ke"] * rt, rel=1e-13, abs=1e-300)
        ic = np.concatenate([np.asarray(f.y_ini(k, tau, p)) for f in fs.user if f.num_equations])
        np.testing.assert_allclose([e.eval(env) for e in fs.ics], ic, rtol=1e-13)


def test_generated_code_is_deterministic():
    a = Model(user_species=(EX.SelfInteractingDR,), **SMALL).fluids
    b = Model(user_species=(EX.SelfInteractingDR,), **SMALL).fluids
    assert a.signature == b.signature and a.dev_src == b.dev_src
    assert abi.host_lib(a.host_path).abcmb_host_fluid_signature().decode() == a.signature


def _split_models():
    lcdm = Model(**SMALL)
    clone = Model(user_species=(CDMClone,), **SMALL)
    pa = lcdm.derived(dict(P0))
    pb = clone.derived(dict(P0, omega_cdm=0.08, omega_x=0.04))
    return lcdm, clone, pa, pb


def test_split_cdm_derived_parameters_and_background():
    """Cold dark matter split between the standard species and a user clone: the
    closure (abcmb/main.py:567-594) and the host producer's tables are those of
    the unsplit model, x_e up to the two-photon stage (HyRex's two-photon fit
    reads params['omega_cdm'] itself, abcmb/hyrex/hydrogen.py:538-575, so from
    there on the split model's history differs, in the reference too)."""
    lcdm, clone, pa, pb = _split_models()
    for key in ("omega_m", "omega_r", "omega_Lambda", "R_b", "R_nu", "om"):
        assert pb[key] == pytest.approx(pa[key], rel=1e-14)
    ba, bb = lcdm.prepare(pa).bg, clone.prepare(pb).bg
    assert bb.tau0 == pytest.approx(ba.tau0, rel=1e-12)
    np.testing.assert_allclose(bb.tau_tab, ba.tau_tab, rtol=1e-12)
    assert np.array_equal(ba.rec_switch[:4], bb.rec_switch[:4])
    n2g = ba.rec_switch[3]
    np.testing.assert_allclose(bb.xe_tab[:n2g], ba.xe_tab[:n2g], rtol=1e-10, atol=1e-14)


def test_oracle_fluid_callbacks_restate_the_standard_cdm():
    """The oracle's user-fluid hooks (metric sums, rows, initial conditions) with a
    clone of cold dark matter: every standard row's derivative is the unsplit
    model's and the clone's row is delta_c's, at random states (the clone's
    entry set to delta_c).  Both read the unsplit model's background tables
    (the split one's recombination differs through HyRex's omega_cdm-dependent
    two-photon fit)."""
    import ctypes as C
    lcdm, clone, pa, pb = _split_models()
    plan_a = lcdm.prepare(pa)
    cb = abi.AbcmbCosmo.from_buffer_copy(plan_a.bg.cosmo)
    cb.omega_cdm = pb["omega_cdm"]
    cb.fluid_par[:] = clone.fluids.fluid_par(pb)

    class _B:
        pass
    plan_b = _B()
    plan_b.bg = _B()
    plan_b.bg.cosmo = cb
    n = lcdm.settings.n
    assert clone.settings.n == n + 1 and clone.fluids.first_row == n
    rng = np.random.default_rng(5)
    for _ in range(6):
        k, lna = 10 ** rng.uniform(-4, 0), rng.uniform(-14, 0)
        y = rng.standard_normal(n)
        ya = O.rhs(plan_a.bg.cosmo, lcdm.settings, k, lna, y)
        ia = O.ic(plan_a.bg.cosmo, lcdm.settings, k, lna)
        with O.fluids(clone, pb):
            yb = O.rhs(plan_b.bg.cosmo, clone.settings, k, lna, np.append(y, y[1]))
            ib = O.ic(plan_b.bg.cosmo, clone.settings, k, lna)
        np.testing.assert_allclose(yb[:n], ya, rtol=1e-11, atol=1e-11 * np.abs(ya).max())
        assert yb[n] == pytest.approx(ya[1], rel=1e-11)
        np.testing.assert_allclose(ib[:n], ia, rtol=1e-12, atol=1e-300)
        assert ib[n] == pytest.approx(ia[1], rel=1e-12)


def test_w0wa_at_minus_one_is_lcdm():
    """w(a) = -1 (w0 = -1, wa = 0): the fluid's density is the cosmological
    constant's, so the derived parameters, the host tables and the oracle's
    perturbation table are LCDM's."""
    lcdm, de = Model(**SMALL), _w0wa()
    assert de.species_off == abi.NO_LAMBDA and de.n_fluid == 0
    pa = lcdm.derived(dict(P0))
    pb = de.derived(dict(P0, w0DE=-1.0, waDE=0.0))
    assert pb["omega_Lambda"] == pytest.approx(pa["omega_Lambda"], rel=1e-14)
    plan_a, plan_b = lcdm.prepare(pa), de.prepare(pb)
    assert plan_b.bg.tau0 == pytest.approx(plan_a.bg.tau0, rel=1e-13)
    np.testing.assert_allclose(plan_b.bg.tau_tab, plan_a.bg.tau_tab, rtol=1e-13)
    from oracle import pipeline
    ks = [0, 3, 5]
    ra = pipeline.run(lcdm, plan_a, k_subset=ks)
    rb = pipeline.run(de, plan_b, k_subset=ks)
    assert np.array_equal(ra["steps"], rb["steps"])
    for f in range(abi.N_PT_FIELDS):
        a, b = ra["pt"][f][:, ks], rb["pt"][f][:, ks]
        np.testing.assert_allclose(b, a, rtol=0, atol=1e-10 * max(np.abs(a).max(), 1e-300))


def test_w0wa_moves_the_background():
    de = _w0wa()
    pa = de.derived(dict(P0, w0DE=-1.0, waDE=0.0))
    pb = de.derived(dict(P0, w0DE=-0.8, waDE=-0.6))
    ta, tb = de.prepare(pa).bg.tau0, de.prepare(pb).bg.tau0
    assert abs(tb / ta - 1) > 1e-3


class _Nonlinear(SP.StandardFluid):
    name = "Nonlinear"
    num_equations = 1

    def rho(self, lna, p):
        return 1e-10 / SP.jnp.exp(lna) ** 3

    def P(self, lna, p):
        return 0.0

    def y_ini(self, k, tau, p):
        return SP.jnp.array([0.0])

    def y_prime(self, k, lna, hp, ep, y, args):
        return SP.jnp.array([y[self.first_idx] * y[self.first_idx]])


class _ToBaryons(_Nonlinear):
    name = "ToBaryons"

    def y_prime(self, k, lna, hp, ep, y, args):
        return SP.jnp.array([y[3] - y[self.first_idx]])


class _TimeMetric(_Nonlinear):
    name = "TimeMetric"

    def y_prime(self, k, lna, hp, ep, y, args):
        return SP.jnp.array([-SP.jnp.exp(lna) * hp])


@pytest.mark.parametrize("cls,match", [(_Nonlinear, "linear"), (_ToBaryons, "standard species"),
                                       (_TimeMetric, "coefficient of h")])
def test_refused_forms(cls, match):
    with pytest.raises(NotImplementedError, match=match):
        Model(user_species=(cls,), **SMALL)


def test_user_species_must_be_classes():
    with pytest.raises(TypeError):
        Model(user_species=(EX.SelfInteractingDR(0, {}),), **SMALL)
    with pytest.raises(NotImplementedError, match="Photon"):
        Model(user_species=(SP.Baryon, SP.MasslessNeutrino), use_LCDM_species=False, **SMALL)


def test_fixed_fluid_parameters_reach_every_parameter_set():
    from abcmb_amd import likelihood as LK
    m = Model(user_species=(EX.SelfInteractingDR,), **SMALL).fix_params(N_idr=0.3)
    p = m.derived(LK.theta_to_params(LK.FIDUCIAL))
    assert p["N_idr"] == 0.3
    assert m.derived(dict(LK.theta_to_params(LK.FIDUCIAL), N_idr=0.1))["N_idr"] == 0.1
    assert m.fluids.fluid_par(p)[0] == 0.3
    with pytest.raises(NotImplementedError, match="forward-mode"):
        LK.adjoint_forward(m, LK.FIDUCIAL[None])


@pytest.mark.parametrize("which", ["w0wa", "nadm"])
def test_forward_host_derivatives_through_fluid_code(which):
    """The host producer's forward mode through the generated fluid code: the
    tangent seeds of tau(ln a), x_e and the LOS background vectors within 1e-6 of
    each derivative's maximum of central differences of frozen-grid producer
    runs at a relative step of 1e-5 (their truncation is ~1e-10 there), with
    the fluids' parameters held; w0wa's density follows omega_Lambda, so h moves
    it, NADM's dark matter and radiation carry omega_cdm's and T_CMB's."""
    from abcmb_amd import likelihood as LK
    from abcmb_amd.model import tangent_seeds
    m = (_w0wa() if which == "w0wa" else _nadm()).fix_params(**EX.PARAMS[which])
    th = LK.FIDUCIAL.copy()
    base = m.prepare(m.derived(LK.theta_to_params(th)))

    def plans(rel, forward):
        st = rel * np.abs(th)
        q = LK._pm_params(m, th, st)
        sd = [st[p] for p in LK.ODE_DIRS]
        pm = m.tangent_plans(base, q, sd) if forward else [m.prepare(x, freeze=base) for x in q]
        return pm, sd

    fwd, sf = plans(1e-4, True)
    fd, s5 = plans(1e-5, False)
    for get in (lambda p: p.bg.tau_tab, lambda p: p.bg.xe_tab, lambda p: p.bgv):
        a = np.stack([(get(fwd[2 * d]) - get(fwd[2 * d + 1])) / (2 * sf[d]) for d in range(len(sf))])
        b = np.stack([(get(fd[2 * d]) - get(fd[2 * d + 1])) / (2 * s5[d]) for d in range(len(s5))])
        a, b = a.reshape(a.shape[0], -1), b.reshape(b.shape[0], -1)
        sc = np.abs(b).max(axis=1)
        for d in range(len(sf)):
            if sc[d] > 0:
                assert np.abs(a[d] - b[d]).max() < 1e-6 * sc[d], (which, d)
    # the seeds' fluid-parameter slots are zero along the standard directions
    seeds = tangent_seeds(fwd[0::2], fwd[1::2], sf)
    assert np.all(seeds.par[:, 31:] == 0.0)


def test_cl_jacobian_refuses_bad_directions():
    m = _w0wa().fix_params(**EX.PARAMS["w0wa"])
    with pytest.raises(ValueError):
        m.cl_jacobian(P0, [])
    with pytest.raises(ValueError):
        m.cl_jacobian(P0, ["h", "omega_b", "omega_cdm", "w0DE", "waDE", "tau_reion"])
    with pytest.raises(ValueError):
        m.cl_jacobian(P0, ["n_s"])
    with pytest.raises(KeyError):
        m.cl_jacobian(P0, ["no_such_parameter"])
