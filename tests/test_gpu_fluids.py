"""GPU path with user-defined fluids against the CPU oracle.

The three models of the reference's fluid tutorial (abcmb_amd.examples:
self-interacting dark radiation on LCDM; interacting dark matter and dark
radiation in place of cold dark matter; w0wa dark energy in place of the
cosmological constant), with the tutorial's parameters.  The GPU runs the
kernels generated from the fluids' methods (abcmb_amd.fluidgen, built ahead
by ``__graft_entry__.build``); the oracle calls the same methods with numbers
through its fluid callbacks.  Tolerances as the standard parity tests: the
same step counts on at least 95% of the k-modes, table fields (and the user
rows) within 1e-4 of each field's maximum (1e-3 on modes whose step counts
differ), unlensed C_l within 1e-6.  Also:
w0wa at w = -1 gives LCDM's spectra on the GPU, and the tangent-linear kernel
with fluid rows keeps the primal kernel's values and matches central
differences of GPU runs.
"""
import numpy as np
import pytest
from _parity import assert_tables_match

from abcmb_amd import examples as EX

pytestmark = pytest.mark.gpu

KK = np.geomspace(1e-4, 0.25, 40)
SPECS = dict(l_max=1000, l_max_g=12, l_max_pol_g=12, l_max_massless_nu=12, max_steps_PE=6000,
             k_axis_perturbations=KK, build_fluids=False)
P0 = dict(h=0.6736, omega_b=0.02237, omega_cdm=0.12, A_s=2.1e-9, n_s=0.9649, tau_reion=0.0544, Neff=3.044)

PARAMS = EX.PARAMS


def _say(msg):
    print(msg, flush=True)


def _model(which, **kw):
    from abcmb_amd.model import Model
    return Model(**EX.model_kwargs(which), **dict(SPECS, **kw))


def _same_steps(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.mean(a == b) >= 0.95 and np.max(np.abs(a - b) / b) <= 0.05


@pytest.mark.parametrize("which", ["sidr", "nadm", "w0wa"])
def test_fluid_model_matches_oracle(which):
    from oracle import pipeline
    m = _model(which)
    plan = m.prepare(m.derived(dict(P0, **PARAMS[which])))
    _say(f"case: fluids {which} on the GPU ({m.fluids.signature})")
    g = m.execute([plan])[0]
    _say(f"case: fluids {which} in the oracle")
    o = pipeline.run(m, plan)
    assert _same_steps(g.steps, o["steps"])
    assert_tables_match(g.PT.table.cpu().numpy(), o["pt"], g.steps, o["steps"], scale_axes=(1, 2))
    if m.n_fluid:
        assert_tables_match(g.PT.pt_fluid.cpu().numpy(), o["pt_fluid"], g.steps, o["steps"], scale_axes=(1, 2))
        for f in m.fluids.user:
            for key, arr in g.PT.species_perturbations[f.name].items():
                assert arr.shape == (m.n_save, KK.size), (f.name, key)
    clg, clo = g.cl_unlensed.cpu().numpy(), o["cl_full"]
    assert np.max(np.abs(clg[0] / clo[0] - 1)) < 1e-6
    assert np.max(np.abs(clg[2] / clo[2] - 1)) < 1e-6
    assert np.max(np.abs(clg[1] - clo[1])) < 1e-6 * np.max(np.abs(clo[1]))


def test_fluids_change_the_spectra():
    """The fluids act: SIDR at N_idr = 0.3 moves TT by more than 1e-3 against LCDM
    on the same k-grid (the tutorial's peak shift)."""
    from abcmb_amd.model import Model
    m, lcdm = _model("sidr"), Model(**SPECS)
    ga = lcdm.execute([lcdm.prepare(lcdm.derived(dict(P0)))])[0].cl_unlensed.cpu().numpy()
    gb = m.execute([m.prepare(m.derived(dict(P0, N_idr=0.3)))])[0].cl_unlensed.cpu().numpy()
    assert np.max(np.abs(gb[0] / ga[0] - 1)) > 1e-3


def test_w0wa_at_minus_one_is_lcdm_on_gpu():
    """Same step counts; tables within 5e-8 and spectra within 2e-7 of their
    maxima: the fluid build is another compilation of the kernels (its density
    sum has one more term, in another order, and multiply-adds contract
    differently), so the two integrations differ at rounding level, grown
    through the ODE, as the tangent-linear and value-only kernels do (measured
    1.5e-8 on the tables, up to 8.6e-8 of a spectrum's maximum)."""
    from abcmb_amd.model import Model
    m, lcdm = _model("w0wa"), Model(**SPECS)
    _say("case: w0wa(-1, 0) vs LCDM on the GPU")
    ga = lcdm.execute([lcdm.prepare(lcdm.derived(dict(P0)))])[0]
    gb = m.execute([m.prepare(m.derived(dict(P0, w0DE=-1.0, waDE=0.0)))])[0]
    assert np.array_equal(np.asarray(ga.steps), np.asarray(gb.steps))
    ta, tb = ga.PT.table.cpu().numpy(), gb.PT.table.cpu().numpy()
    scale = np.abs(ta).max(axis=(1, 2), keepdims=True) + 1e-300
    assert np.max(np.abs(tb - ta) / scale) < 5e-8
    ca, cb = ga.cl_unlensed.cpu().numpy(), gb.cl_unlensed.cpu().numpy()
    for x in range(3):
        assert np.max(np.abs(cb[x] - ca[x])) < 2e-7 * np.max(np.abs(ca[x]))


@pytest.mark.parametrize("which", ["sidr", "nadm"])
def test_tangent_tables_with_fluids_match_oracle(which):
    """The tangent-linear kernel built with the user fluids' rows against the
    oracle's tangent-linear integration (same step sequence, stage tangents
    solved to convergence, the right-hand side's parameter derivatives by
    central differences of the fluids' methods between neighbouring
    cosmologies), along the five ODE directions, as the standard build is
    checked: values within the Newton tolerance of the primal kernel's, tangents
    within 1e-4 of each field's maximum on the k-modes whose step counts agree."""
    from abcmb_amd import likelihood as LK
    from abcmb_amd.model import tangent_seeds
    from oracle import pipeline
    kk = np.geomspace(1e-4, 0.3, 8)
    m = _model(which, l_max=500, k_axis_perturbations=kk)
    th = LK.FIDUCIAL.copy()
    extra = PARAMS[which]
    par = lambda t: dict(LK.theta_to_params(t), **extra)
    base = m.prepare(m.derived(par(th)))
    steps = LK.FD_REL_STEP * np.abs(th)
    plus, minus = [], []
    for d in LK.ODE_DIRS:
        for sgn, lst in ((1, plus), (-1, minus)):
            t = th.copy()
            t[d] += sgn * steps[d]
            lst.append(m.prepare(m.derived(par(t)), freeze=base))
    sd = [steps[d] for d in LK.ODE_DIRS]
    _say(f"case: tangent tables with fluids {which} vs oracle")
    outs, dpts = m.execute_tangent([base], [tangent_seeds(plus, minus, sd)])
    prim = m.execute([base])[0]
    pt_t, pt_p = outs[0].PT.table.cpu().numpy(), prim.PT.table.cpu().numpy()
    sp = np.asarray(outs[0].steps) == np.asarray(prim.steps)
    assert np.mean(sp) >= 0.75
    scale = np.abs(pt_p).max(axis=(1, 2), keepdims=True) + 1e-300
    dev = (np.abs(pt_t - pt_p) / scale).max(axis=(0, 1))
    _say(f"tangent-kernel values vs primal per k: {np.array2string(dev, precision=1)}")
    assert np.max(dev[sp]) < 5e-6
    pt_o, dpt_o, steps_o = pipeline.tangent_tables(m, base, plus, minus, sd, np.arange(kk.size))
    same = np.asarray(outs[0].steps) == steps_o
    assert np.mean(same) >= 0.75
    dg = dpts[0].cpu().numpy()[..., same]
    do = dpt_o[..., same]
    sc = np.abs(do).max(axis=(2, 3), keepdims=True) + 1e-300
    err = np.abs(dg - do) / sc
    _say(f"max tangent error vs oracle {err.max():.2e}")
    assert err.max() < 1e-4


def test_forward_mode_gradients_with_fluids():
    """Likelihood gradients of a model with SIDR (its parameter held by
    fix_params) from the tangent-linear path against central finite differences
    of full evaluations at 100x the reference's ODE tolerances, where the two
    converge together (5e-4 of the gradient norm for the standard build)."""
    from abcmb_amd import likelihood as LK
    tol = 1e-2
    m = _model("sidr", lensing=True, l_max=1000, k_axis_perturbations=None, max_steps_PE=30000,
               rtol_small_k_PE=1e-5 * tol, rtol_large_k_PE=1e-4 * tol, atol_small_k_PE=1e-10 * tol,
               atol_large_k_PE=1e-6 * tol).fix_params(**PARAMS["sidr"])
    thetas = LK.latin_hypercube(1, seed=2)
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=1, seed=1)
    _say("case: tangent vs finite-difference gradients with SIDR")
    Lf, gf = LK.fd_gradients(m, like, thetas, batched=True)
    Lt, gt = LK.tangent_gradients(m, like, thetas, downstream="fd")
    assert np.allclose(Lf, Lt, rtol=1e-6, atol=0)
    scale = np.linalg.norm(gf * thetas, axis=1, keepdims=True) / np.abs(thetas)
    err = np.abs(gt - gf) / scale
    _say(f"tangent vs FD, relative to the norm: {err.max():.2e}")
    assert err.max() < 5e-4


def test_cl_jacobian_along_a_fluid_parameter():
    """Derivatives along a user fluid's own parameter (SIDR's N_idr, which
    moves radiation from the free-streaming neutrinos into the fluid at fixed
    N_eff): the tangent-linear tables against the oracle's tangent-linear
    integration with the same seeds (its right-hand-side derivative by central
    differences of the fluid's methods between the neighbouring cosmologies),
    within 1e-4 of each field's maximum on the k-modes whose step counts agree
    (measured on MI355X: 2.6e-5);
    Model.cl_jacobian's value part equals the plain evaluation's to 1e-5 of each
    spectrum's maximum (the tangent-linear LOS evaluates the Bessel functions by
    interpolated segments and asymptotic forms; measured 3.4e-6), and its
    N_idr column is the downstream of those tables.  (Central differences of
    full evaluations cannot check it: d C_l / d N_idr is ~1e-4 of C_l per unit,
    below their step-size-control noise even at 100x tighter tolerances.)"""
    from abcmb_amd import likelihood as LK
    from abcmb_amd.model import tangent_seeds
    from oracle import pipeline
    kk = np.geomspace(1e-4, 0.3, 8)
    m = _model("sidr", l_max=500, k_axis_perturbations=kk).fix_params(**PARAMS["sidr"])
    full = m.derived(P0)
    base = m.prepare(full)
    v = float(full["N_idr"])
    h = LK.FD_REL_STEP * v
    pm = m.tangent_plans(base, [m.derived(dict(P0, N_idr=v + s * h)) for s in (1, -1)], [h])
    plus, minus = pm[0::2], pm[1::2]
    _say("case: tangent tables along N_idr vs oracle")
    outs, dpts = m.execute_tangent([base], [tangent_seeds(plus, minus, [h])])
    pt_o, dpt_o, steps_o = pipeline.tangent_tables(m, base, plus, minus, [h], np.arange(kk.size))
    same = np.asarray(outs[0].steps) == steps_o
    assert np.mean(same) >= 0.75
    dg = dpts[0].cpu().numpy()[:1][..., same]
    do = dpt_o[:1][..., same]
    sc = np.abs(do).max(axis=(2, 3), keepdims=True) + 1e-300
    err = np.abs(dg - do) / sc
    _say(f"max tangent error along N_idr vs oracle {err.max():.2e}")
    assert err.max() < 1e-4
    J = m.cl_jacobian(P0, ["N_idr"])
    n_l = J["cl"].shape[1]
    c0 = m.execute([base])[0].cl_unlensed.cpu().numpy()[:, :n_l]
    assert np.max(np.abs(J["cl"] - c0) / np.abs(c0).max(axis=1, keepdims=True)) < 1e-5
    assert np.all(np.isfinite(J["N_idr"])) and np.abs(J["N_idr"][0]).max() > 0


def test_autograd_spectra_along_fluid_parameters():
    """autograd.spectra: a loss of SIDR's spectra back-propagates to N_idr and
    omega_cdm as the contraction of Model.cl_jacobian's columns (1e-12)."""
    import torch
    from abcmb_amd.autograd import spectra
    m = _model("sidr", l_max=500, k_axis_perturbations=np.geomspace(1e-4, 0.3, 8)).fix_params(**PARAMS["sidr"])
    names = ["N_idr", "omega_cdm"]
    vals = torch.tensor([0.3, 0.12], dtype=torch.float64, requires_grad=True)
    _say("case: autograd.spectra along N_idr and omega_cdm")
    cl = spectra(m, vals, names, P0)
    w = torch.linspace(0.0, 1.0, cl.shape[1], dtype=torch.float64)
    (cl[0].cpu() * w).sum().backward()
    J = m.cl_jacobian(dict(P0, N_idr=0.3, omega_cdm=0.12), names)
    ref = np.array([(J[n][0] * w.numpy()).sum() for n in names])
    assert np.allclose(vals.grad.numpy(), ref, rtol=1e-12, atol=0)
    assert np.array_equal(cl.detach().cpu().numpy(), J["cl"])
