"""GPU tangent-linear path through the C ABI against the CPU oracle.

  - perturbation-table tangents of the HIP kernel against the oracle's
    tangent-linear integration (same step sequence, stage tangents solved to
    convergence): within 1e-4 of each field's maximum on the k-modes whose
    step counts agree (the two step sequences can differ by a rounding-level
    accept/reject flip, like the primal's); the kernel's values equal the
    primal kernel's to 5e-8 of each field's maximum on most of those k-modes
    (the compiler contracts multiply-adds differently when the products also
    feed the tangents: measured 1-2e-8) and within the Newton tolerance (5e-6)
    on the others, where a chord-convergence decision flips between the two
    kernels (measured 1.2e-6 on one k-mode);
  - likelihood gradients from ``tangent_gradients`` against the oracle's
    gradient built the same way on the CPU (oracle tangent tables; oracle
    sources/LOS/lensing differenced along each direction, or the unlensed C_l
    moved along branch-frozen oracle derivatives and the lensing
    differenced): 3e-4 of the gradient norm for omega_cdm, whose distance
    derivative the 24-k-mode mirror pins least tightly (measured 2.6e-5 to
    1.7e-4 over the cases), and 3e-5 (downstream differenced) or 1e-4 (the
    tangent-linear LOS) for the other six, lensed to l = 1000 and, downstream
    differenced, to l = 2500;
  - at 100x tighter ODE tolerances the tangent gradient (downstream
    differenced, as the finite differences do) and central finite differences
    of full evaluations agree to 5e-4 of the gradient norm (at the default
    tolerances their difference is the finite differences' step-size noise, a
    few 1e-3);
  - the host producer's forward-mode seeds give the gradients of frozen-grid
    producer differences to 1e-5 of the norm (measured 2.7e-8);
  - the tangent-linear LOS is the pointwise derivative: branch-frozen oracle
    differences converge to it as the step shrinks;
  - chunks enqueued ahead of the GPU (no host wait between them) are
    deterministic and give a single chunk's likelihoods and gradients to
    rounding.
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pm(m, th, base, k=None, rel=LK.FD_REL_STEP):
    """Plans at theta +- step along the ODE directions, made as the gradient
    pipeline makes them (the host producer's forward-mode derivatives, or
    frozen-grid producer runs for a model with user fluids)."""
    steps = rel * np.abs(th)
    q = LK._pm_params(m, th, steps)
    sd = [steps[p] for p in LK.ODE_DIRS]
    pm = m.tangent_plans(base, q, sd) if LK._forward_host(m) else [m.prepare(x, freeze=base, k=k) for x in q]
    return pm[0::2], pm[1::2], sd


def _norm_err(g, ref, theta):
    """per-parameter |g - ref| relative to the gradient norm in units of each parameter."""
    scale = np.linalg.norm(ref * theta, axis=1, keepdims=True) / np.abs(theta)
    return np.abs(g - ref) / scale


@pytest.mark.parametrize("massive", [False, True])
def test_tangent_tables_match_oracle(massive):
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    from oracle import pipeline
    kk = np.geomspace(1e-4, 0.3, 8)
    m = Model(user_species=(MassiveNeutrino,) if massive else (), l_max=500, k_axis_perturbations=kk)
    th = LK.FIDUCIAL.copy()
    base = m.prepare(m.derived(LK.theta_to_params(th)))
    plus, minus, sd = _pm(m, th, base)
    _say(f"case: tangent tables vs oracle, massive={massive}")
    outs, dpts = m.execute_tangent([base], [tangent_seeds(plus, minus, sd)])
    prim = m.execute([base])[0]
    pt_t, pt_p = outs[0].PT.table.cpu().numpy(), prim.PT.table.cpu().numpy()
    sp = outs[0].steps == prim.steps
    assert np.mean(sp) >= 0.75
    scale = np.abs(pt_p).max(axis=(1, 2), keepdims=True) + 1e-300
    dev = (np.abs(pt_t - pt_p) / scale).max(axis=(0, 1))          # per k-mode
    _say(f"tangent-kernel values vs value-only kernel per k: {np.array2string(dev, precision=1)}")
    # to rounding on most k-modes; where a chord-convergence test flips between the
    # two kernels (their multiply-adds round differently), a stage moves by up to the
    # Newton tolerance, kappa * rtol ~ 1e-6 of the state
    assert np.mean(dev[sp] < 5e-8) >= 0.6
    assert np.max(dev[sp]) < 5e-6
    pt_o, dpt_o, steps_o = pipeline.tangent_tables(m, base, plus, minus, sd, np.arange(kk.size))
    same = outs[0].steps == steps_o
    assert np.mean(same) >= 0.75
    dg = dpts[0].cpu().numpy()[..., same]
    do = dpt_o[..., same]
    scale = np.abs(do).max(axis=(2, 3), keepdims=True) + 1e-300
    err = np.abs(dg - do) / scale
    _say(f"max tangent error {err.max():.2e}")
    assert err.max() < 1e-4


@pytest.mark.parametrize("downstream,l_max", [("fd", 1000), ("tangent", 1000), ("fd", 2500)])
def test_tangent_gradient_matches_oracle(downstream, l_max):
    """Both directional downstreams against the oracle built the same way: "fd"
    differences the oracle's sources / LOS / lensing along each direction;
    "tangent" moves the unlensed C_l along their pointwise derivatives and
    differences the lensing, as ``_directional_losses_linear`` does on the GPU;
    the oracle's pointwise derivatives are branch-frozen differences at the
    step where they are closest to the limit (1e-8 relative for omega_b,
    omega_cdm, h, whose distances move the Bessel arguments across linearly
    interpolated table segments; 1e-6 for tau_reion and m_nu, whose small
    derivatives need a larger step against rounding).  Bounds: 3e-4 of the
    gradient norm for omega_cdm (measured 2.6e-5 and 1.1e-4 for "fd" at l <=
    1000 and 2500, 1.7e-4 for "tangent": on 24 k-modes the oracle's
    branch-frozen differences of the TE derivative along omega_cdm still move
    by ~5e-5 of its maximum between steps of 1e-7 and 1e-8, and the GPU's and
    the oracle's chord iterations can stop on either side of a convergence
    threshold in a few stages, which moves the table by up to the Newton
    tolerance); 3e-5 ("fd") and 1e-4 ("tangent") for the other six (measured
    1.4e-5 and 3e-5)."""
    from abcmb_amd.model import Model, MassiveNeutrino
    from abcmb_amd.likelihood import with_primordial
    from oracle import pipeline
    kk = np.geomspace(1e-4, 0.6, 24)
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=l_max, k_axis_perturbations=kk)
    thetas = LK.latin_hypercube(2, seed=3)
    base0 = m.prepare(m.derived(LK.theta_to_params(thetas[0])))
    g0 = m.execute([base0])[0]
    like = LK.LensedClLikelihood(m, [g0.ClTT.cpu().numpy(), g0.ClTE.cpu().numpy(), g0.ClEE.cpu().numpy()],
                                 n_mocks=2, seed=5)
    _say(f"case: tangent gradients ({downstream} downstream), 2 cosmologies, lensed l<={l_max}, 24 k")
    Lg, gg = LK.tangent_gradients(m, like, thetas, downstream=downstream)
    # L(theta) -- in "tangent" mode lensed from the tangent-linear LOS's value --
    # against a plain evaluation: the tables of the tangent kernel differ from
    # the value-only kernel's by up to 5e-8 of each field's maximum, which
    # moves L by ~1e-7 (measured 1.4e-7)
    Lplain = np.array([float(like.loglike(m.execute([m.prepare(m.derived(LK.theta_to_params(th)))]), [c])[0])
                       for c, th in enumerate(thetas)])
    _say(f"L(theta) vs plain evaluation: {np.max(np.abs(Lg / Lplain - 1)):.1e}")
    assert np.allclose(Lg, Lplain, rtol=1e-6, atol=0)
    mock = like.mock.cpu().numpy()
    ivar = like.inv_var.cpu().numpy()
    n_l = m.ells.size

    def loglike(out, slot):
        cl = out["lensed"][:3, :n_l]
        return -0.5 * np.sum((cl - mock[slot]) ** 2 * ivar[slot])

    go = np.zeros_like(gg)
    for c, th in enumerate(thetas):
        base = m.prepare(m.derived(LK.theta_to_params(th)))
        plus, minus, sd = _pm(m, th, base)
        pt, dpt, _ = pipeline.tangent_tables(m, base, plus, minus, sd, np.arange(kk.size))
        steps = LK.FD_REL_STEP * np.abs(th)
        if downstream == "tangent":
            cl0 = pipeline.downstream(m, base, pt, lensing=False)["cl_full"]
            chi_b = base.bg.tau0 - base.bgv[0]
            small = {rel: _pm(m, th, base, rel=rel) for rel in (1e-8, 1e-6)}

        def moved(p, sgn, plan, tab):
            """loglike at theta + sgn h e_p: the lensing of plan / tab applied to the
            unlensed C_l differenced ("fd") or moved along the derivative ("tangent")"""
            if downstream == "fd":
                return loglike(pipeline.downstream(m, plan, tab), c)
            return loglike(pipeline.downstream(m, plan, tab, cl_full=cl0 + sgn * steps[p] * dcl[p]), c)

        dcl = {}
        for p in range(7):
            if downstream == "tangent":
                if p in LK.ODE_DIRS:
                    d = LK.ODE_DIRS.index(p)
                    plus_s, minus_s, sd_s = small[1e-8 if p in (0, 1, 2) else 1e-6]
                    cp = pipeline.downstream(m, plus_s[d], pt + sd_s[d] * dpt[d], lensing=False, chi_branch=chi_b)
                    cm = pipeline.downstream(m, minus_s[d], pt - sd_s[d] * dpt[d], lensing=False, chi_branch=chi_b)
                    dcl[p] = (cp["cl_full"] - cm["cl_full"]) / (2 * sd_s[d])
                else:
                    cc = []
                    for sgn in (+1, -1):
                        t = th.copy()
                        t[p] += sgn * steps[p]
                        q = LK.theta_to_params(t)
                        cc.append(pipeline.downstream(m, with_primordial(base, q["A_s"], q["n_s"]), pt,
                                                      lensing=False)["cl_full"])
                    dcl[p] = (cc[0] - cc[1]) / (2 * steps[p])
            L = []
            for sgn in (+1, -1):
                if p in LK.ODE_DIRS:
                    d = LK.ODE_DIRS.index(p)
                    plan = (plus if sgn > 0 else minus)[d]
                    tab = pt + sgn * steps[p] * dpt[d]
                else:
                    t = th.copy()
                    t[p] += sgn * steps[p]
                    q = LK.theta_to_params(t)
                    plan, tab = with_primordial(base, q["A_s"], q["n_s"]), pt
                L.append(moved(p, sgn, plan, tab))
            go[c, p] = (L[0] - L[1]) / (2 * steps[p])
    err = _norm_err(gg, go, thetas)
    _say(f"gradient vs oracle, relative to the norm: {err.max():.2e}")
    _say("per parameter: " + " ".join(f"{n} {e:.1e}" for n, e in zip(LK.NAMES, err.max(axis=0))))
    assert err.max() < 3e-4
    assert np.delete(err, LK.NAMES.index("omega_cdm"), axis=1).max() < (3e-5 if downstream == "fd" else 1e-4)


def test_forward_host_seeds_match_frozen_grid_differences():
    """The tangent gradients with the host producer's forward-mode seeds
    (the default) against the same gradients with the seeds from central
    differences of frozen-grid producer runs: the two seed sets differ by the
    differences' truncation and rounding (~1e-7 of each seed's maximum; up to
    2e-4 of dT_m/dh's small maximum), so the gradients agree to 1e-5 of the norm
    (measured on MI355X: 2.7e-8) and L(theta) to rounding."""
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=1000,
              k_axis_perturbations=np.geomspace(1e-4, 0.6, 24))
    thetas = LK.latin_hypercube(2, seed=7)
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=2, seed=3)
    _say("case: tangent gradients, forward-mode vs frozen-grid host seeds")
    assert m.host_tangents == "forward"
    Lf, gf = LK.tangent_gradients(m, like, thetas)
    m.host_tangents = "fd"
    Ld, gd = LK.tangent_gradients(m, like, thetas)
    err = _norm_err(gf, gd, thetas)
    _say(f"forward vs frozen-grid seeds, relative to the norm: {err.max():.2e}")
    assert np.allclose(Lf, Ld, rtol=1e-10, atol=0)
    assert err.max() < 1e-5


def test_cl_jacobian_equals_the_gradient_pipeline():
    """Model.cl_jacobian along the five parameters that reach the ODE builds the
    same plans, seeds, tangent tables and lensing differences as
    tangent_gradients: its lensed derivatives equal the pipeline's Jacobian
    (``jac_out``) to 1e-9 of each spectrum's largest derivative (measured on
    MI355X: equal bit for bit)."""
    import torch
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=1000,
              k_axis_perturbations=np.geomspace(1e-4, 0.6, 24))
    th = LK.latin_hypercube(1, seed=11)
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=1, seed=2)
    _say("case: cl_jacobian vs the gradient pipeline's Jacobian")
    jac = []
    LK.tangent_gradients(m, like, th, jac_out=jac)
    jac = torch.cat(jac)[0].cpu().numpy()                      # [7][3][n_l]
    names = [("omega_b", 0), ("omega_cdm", 1), ("h", 2), ("tau_reion", 5), ("m_nu_massive", 6)]
    J = m.cl_jacobian(LK.theta_to_params(th[0]), [n for n, _ in names])
    for n, p in names:
        a, b = J["lensed_" + n][:3], jac[p]
        err = np.max(np.abs(a - b) / (np.abs(b).max(axis=1, keepdims=True) + 1e-300))
        _say(f"{n}: {err:.1e}")
        assert err < 1e-9, n


def test_tangent_and_fd_gradients_converge_together():
    from abcmb_amd.model import Model, MassiveNeutrino
    tol = 1e-2
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=30000,
              rtol_small_k_PE=1e-5 * tol, rtol_large_k_PE=1e-4 * tol, atol_small_k_PE=1e-10 * tol,
              atol_large_k_PE=1e-6 * tol)
    thetas = LK.latin_hypercube(2, seed=0)
    kg = LK.uneven_k_grids(2, seed=0)
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=2, seed=1)
    _say("case: tangent vs finite-difference gradients at 100x tighter tolerances")
    Lf, gf = LK.fd_gradients(m, like, thetas, k_grids=kg, batched=True)
    Lt, gt = LK.tangent_gradients(m, like, thetas, k_grids=kg, downstream="fd")
    assert np.allclose(Lf, Lt, rtol=1e-6, atol=0)
    err = _norm_err(gt, gf, thetas)
    _say(f"tangent vs FD, relative to the norm: {err.max():.2e}")
    assert err.max() < 5e-4


def test_tangent_los_matches_finite_differences():
    """``abcmb_los_tangent`` (one tangent-linear LOS pass) against central
    differences of the FP64 LOS at theta +- h with h shrinking: the Bessel
    lookups are differentiated as interpolated segments, so the differences
    converge to the tangent as h -> 0 (h = 1e-6 relative: 1e-4 of each
    derivative's maximum; omega_cdm 1e-3, where table/asymptotic branch
    switches at x_max -- a 4e-5 jump -- enter the differences but not the
    derivative).  The value is the FP64 LOS's to 1e-12."""
    import torch
    from abcmb_amd.model import Model, MassiveNeutrino
    from abcmb_amd.likelihood import with_primordial
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096,
              k_axis_perturbations=np.geomspace(7.05e-6, 0.62, 200))
    thetas = LK.latin_hypercube(1, seed=4)
    rel = 1e-6
    steps = rel * np.abs(thetas)
    _say("case: tangent-linear LOS vs finite differences (h = 1e-6)")
    base_plans, pm, seeds = LK._tangent_inputs(m, thetas, [None], steps)
    outs, dpts = m.execute_tangent(base_plans, seeds)
    pts = [outs[0].PT.table]
    kdev = [outs[0].PT.k_dev]
    plus, minus = [pm[0][0::2]], [pm[0][1::2]]
    sd = [[steps[0, p] for p in LK.ODE_DIRS]]
    cl_t = m.downstream_tangent(base_plans, pts, kdev, dpts, plus, minus, sd).cpu().numpy()[0]
    c0 = outs[0].cl_unlensed.cpu().numpy()
    assert np.max(np.abs(cl_t[0][[0, 2]] / c0[[0, 2]] - 1)) < 1e-12
    for d, p in enumerate(LK.ODE_DIRS):
        vt = [pts[0] + sd[0][d] * dpts[0][d], pts[0] - sd[0][d] * dpts[0][d]]
        ds = m._downstream([plus[0][d], minus[0][d]], vt, kdev * 2)
        fd = (ds["cl_full"][0] - ds["cl_full"][1]).cpu().numpy() / (2 * sd[0][d])
        err = np.max(np.abs(cl_t[1 + d] - fd)) / np.max(np.abs(fd))
        _say(f"{LK.NAMES[p]}: {err:.1e}")
        assert err < (1e-3 if LK.NAMES[p] == "omega_cdm" else 1e-4)
    for p, row in ((3, 1 + len(LK.ODE_DIRS)), (4, 2 + len(LK.ODE_DIRS))):
        q = [LK.theta_to_params(thetas[0] + s * steps[0, p] * np.eye(7)[p]) for s in (1, -1)]
        ds = m._downstream([with_primordial(base_plans[0], x["A_s"], x["n_s"]) for x in q], pts * 2, kdev * 2)
        fd = (ds["cl_full"][0] - ds["cl_full"][1]).cpu().numpy() / (2 * steps[0, p])
        err = np.max(np.abs(cl_t[row] - fd)) / np.max(np.abs(fd))
        _say(f"{LK.NAMES[p]}: {err:.1e}")
        assert err < 1e-4


def test_tangent_los_is_the_pointwise_derivative():
    """``abcmb_los_tangent`` against central differences of the oracle's LOS with
    the Bessel branches (zero / table / asymptotic form) picked at the base
    cosmology's distances (``oracle.los_cl_branch``).  With the branch jumps
    out of the differences (the table and the asymptotic form differ by ~4e-5
    of the function's maximum where they meet, and differences of the
    reference's own branch choice differ from the derivative by up to 1e-3 in
    omega_cdm) they converge to the pointwise derivative -- what autodiff of
    the reference's LOS gives: from h = 1e-6 to 1e-7 relative the distance-
    moving directions (omega_b, omega_cdm, h) come 4x or more closer (measured:
    ~25x, first order in h -- the tables are interpolated linearly, so the
    differences also see the slope changes between segments), and at 1e-7
    every direction agrees to 2e-5 of its maximum.  Sources of the tables at
    theta +- h come from the GPU; the LOS and l-spline of the differences are
    the oracle's."""
    from abcmb_amd.model import Model, MassiveNeutrino
    from oracle import oracle as O
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096,
              k_axis_perturbations=np.geomspace(7.05e-6, 0.62, 200))
    thetas = LK.latin_hypercube(1, seed=4)
    errs = {}
    for rel in (1e-6, 1e-7):
        steps = rel * np.abs(thetas)
        _say(f"case: tangent-linear LOS vs branch-frozen oracle differences (h = {rel:.0e})")
        base_plans, pm, seeds = LK._tangent_inputs(m, thetas, [None], steps)
        outs, dpts = m.execute_tangent(base_plans, seeds)
        pts, kdev = [outs[0].PT.table], [outs[0].PT.k_dev]
        plus, minus = [pm[0][0::2]], [pm[0][1::2]]
        sd = [[steps[0, p] for p in LK.ODE_DIRS]]
        cl_t = m.downstream_tangent(base_plans, pts, kdev, dpts, plus, minus, sd).cpu().numpy()[0]
        base = base_plans[0]
        chi_b = base.bg.tau0 - base.bgv[0]
        for d, p in enumerate(LK.ODE_DIRS):
            pls = [plus[0][d], minus[0][d]]
            vt = [pts[0] + sd[0][d] * dpts[0][d], pts[0] - sd[0][d] * dpts[0][d]]
            src = m._run_sources(pls, vt, kdev * 2).cpu().numpy()
            cl = [O.los_cl_branch(src[i], pl.bgv[0], pl.bg.tau0, m.k_transfer, m.bes_tab, m.bes_rng, m.bes_l,
                                  m.ell_idx, base.bg.cosmo.A_s, base.bg.cosmo.n_s, m.specs["k_pivot"], chi_b)
                  for i, pl in enumerate(pls)]
            fd = ((cl[0] - cl[1]) / (2 * sd[0][d])) @ m.ell_op.T
            errs[rel, p] = np.max(np.abs(cl_t[1 + d] - fd)) / np.max(np.abs(fd))
            _say(f"{LK.NAMES[p]}: {errs[rel, p]:.1e}")
    for p in LK.ODE_DIRS:
        assert errs[1e-7, p] < 2e-5
    for p in (0, 1, 2):
        assert errs[1e-7, p] < 0.25 * errs[1e-6, p]


def test_pipelined_chunks_match_one_launch():
    """Chunks enqueued ahead of the GPU (ODE launch, downstream and losses of
    chunk i + 1 queued behind chunk i's, asynchronous staging copies, one
    wait at the end) are deterministic, and agree with a single chunk to
    rounding: every kernel's result for a cosmology is independent of which
    other cosmologies share its launch; only torch's batched sums in the loss
    may add in another order."""
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=1000)
    thetas = LK.latin_hypercube(7, seed=3)
    kg = LK.uneven_k_grids(7, seed=3, hi=0.3)
    kg = [k[::4] for k in kg]
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=7, seed=1)
    _say("case: pipelined chunks of 1, 1, 2, 3 cosmologies (twice) vs one chunk of 7")
    Lp, gp = LK.tangent_gradients(m, like, thetas, k_grids=kg, chunk=4)
    Lq, gq = LK.tangent_gradients(m, like, thetas, k_grids=kg, chunk=4)
    L1, g1 = LK.tangent_gradients(m, like, thetas, k_grids=kg, chunk=64)
    assert np.array_equal(Lp, Lq) and np.array_equal(gp, gq)
    assert np.all(np.isfinite(gp))
    dl = np.max(np.abs(Lp - L1) / np.abs(L1))
    dg = _norm_err(gp, g1, thetas).max()
    _say(f"pipelined vs one chunk: L {dl:.1e}, gradient {dg:.1e} of the norm")
    assert dl < 1e-13
    assert dg < 1e-8
