"""GPU tests of the workloads at their full sizes, through the C ABI.

  - configuration 2 (one massive neutrino, 500 k-modes, lensed to l = 2500,
    FP64): the whole GPU path against a complete oracle evaluation of the
    same inputs (every k-mode, sources, LOS, C_l^phiphi, lensing): lensed
    TT/EE/BB within 1e-6 relative per l, TE within 1e-6 of max|TE|,
    C_l^phiphi within 1e-6; the same with five momentum bins (state size 136);
  - configuration 3 (64 Latin-hypercube cosmologies, each with its own
    300-800-mode k grid, lensed to l = 2500, FP32 LOS, one ODE launch):
    three of them against complete oracle evaluations within the contract
    (1e-4 on TT/TE/EE, 1e-3 on BB) and four equal bit for bit to their own
    single-cosmology runs;
  - configuration 5 (256 Latin-hypercube cosmologies, own k grids, batched
    tangent-linear gradients): every likelihood and gradient finite, and the
    default chunking (8, 16, 32, then 64 cosmologies per ODE launch) equal to
    a single 256-cosmology launch to rounding.
"""
import numpy as np
import pytest
from _parity import assert_tables_match

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu

P0 = {"h": 0.6736, "omega_b": 0.02237, "omega_cdm": 0.1200, "A_s": 2.1e-9, "n_s": 0.9649, "tau_reion": 0.0544}


def _say(msg):
    print(msg, flush=True)


def _lensed(o):
    return [x.cpu().numpy() for x in (o.ClTT, o.ClTE, o.ClEE, o.ClBB)]


def _errors(g, ref):
    """TT, TE, EE, BB errors: relative per l (TE relative to max|TE|)."""
    n = g[0].size
    e = []
    for x in range(4):
        r = ref[x][:n]
        den = np.max(np.abs(r)) if x == 1 else np.abs(r)
        e.append(float(np.max(np.abs(g[x] - r) / den)))
    return e


def _same_steps(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.mean(a == b) >= 0.95


@pytest.mark.parametrize("bins", [3, 5])
def test_config2_lensed_l2500_matches_oracle(bins):
    from abcmb_amd.model import Model, MassiveNeutrino
    from oracle import pipeline
    m = Model(user_species=(MassiveNeutrino,), massive_nu_bins=bins, lensing=True,
              k_axis_perturbations=np.geomspace(1e-5, 0.5, 500), max_steps_PE=4096)
    assert m.layout.n == (100 if bins == 3 else 136)
    plan = m.prepare(m.derived(dict(P0, N_nu_massive=1, m_nu_massive=0.06)))
    _say(f"case: config2, {bins} momentum bins, 500 k, lensed l<=2500, GPU vs complete oracle")
    g = m.execute([plan])[0]
    o = pipeline.run(m, plan)
    assert np.all(o["status"] == 0)
    assert _same_steps(g.steps, o["steps"])
    assert_tables_match(g.PT.table.cpu().numpy(), o["pt"], g.steps, o["steps"], scale_axes=(1, 2))
    assert np.max(np.abs(g.Clpp.cpu().numpy() / o["clpp"] - 1)) < 1e-6
    e = _errors(_lensed(g), o["lensed"])
    _say(f"lensed vs oracle: TT {e[0]:.1e} TE {e[1]:.1e} EE {e[2]:.1e} BB {e[3]:.1e}")
    assert max(e) < 1e-6


def test_config2_2000_los_samples():
    """Configuration 2's 2000 conformal-time LOS samples (n_lna_save = 2001 saved
    points: the ODE's dense output, sources, LOS and Limber on that grid), FP64 and
    FP32 LOS, against a complete oracle evaluation on the same grid: FP64 within
    1e-6, FP32 within the contract; the spectra move from the 499-sample ones by the
    LOS quadrature's own error (printed)."""
    from abcmb_amd.model import Model, MassiveNeutrino
    from oracle import pipeline
    kw = dict(user_species=(MassiveNeutrino,), lensing=True, k_axis_perturbations=np.geomspace(1e-5, 0.5, 500),
              max_steps_PE=4096)
    m = Model(n_lna_save=2001, **kw)
    m32 = Model(n_lna_save=2001, los_precision="fp32", **kw)
    m500 = Model(**kw)
    p = m.derived(dict(P0, N_nu_massive=1, m_nu_massive=0.06))
    plan = m.prepare(p)
    assert plan.lna_save.size == 2001 and plan.bgv.shape == (6, 2000)
    _say("case: config2, 2000 LOS samples, FP64 and FP32 LOS vs complete oracle")
    g = m.execute([plan])[0]
    g32 = m32.execute([plan])[0]
    o = pipeline.run(m, plan)
    assert np.all(o["status"] == 0)
    e = _errors(_lensed(g), o["lensed"])
    e32 = _errors(_lensed(g32), o["lensed"])
    e500 = _errors(_lensed(m500.execute([m500.prepare(p)])[0]), o["lensed"])
    _say(f"FP64 vs oracle: TT {e[0]:.1e} TE {e[1]:.1e} EE {e[2]:.1e} BB {e[3]:.1e}; FP32: TT {e32[0]:.1e} "
         f"TE {e32[1]:.1e} EE {e32[2]:.1e} BB {e32[3]:.1e}; 499 samples vs 2000: TT {e500[0]:.1e} "
         f"TE {e500[1]:.1e} EE {e500[2]:.1e} BB {e500[3]:.1e}")
    assert max(e) < 1e-6
    assert max(e32[:3]) < 1e-4 and e32[3] < 1e-3


def test_config3_batched_forward():
    from abcmb_amd.model import Model, MassiveNeutrino
    from oracle import pipeline
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096, los_precision="fp32")
    thetas = LK.latin_hypercube(64, seed=2)
    kg = LK.uneven_k_grids(64, seed=2)
    assert min(k.size for k in kg) >= 300 and max(k.size for k in kg) <= 800
    plans = m.prepare_many([m.derived(LK.theta_to_params(t)) for t in thetas], k_list=kg)
    _say("case: config3, 64 cosmologies with their own k grids in one ODE launch, FP32 LOS")
    outs = m.execute(plans)
    for o in outs:
        for x in _lensed(o):
            assert np.all(np.isfinite(x))
        assert np.all(o.ClTT.cpu().numpy() > 0) and np.all(o.ClBB.cpu().numpy()[20:] > 0)
    for c in (0, 23, 47, 63):
        single = m.execute([plans[c]])[0]
        for a, b in zip(_lensed(outs[c]), _lensed(single)):
            assert np.array_equal(a, b)
    m64 = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096)
    for c in (5, 40, 60):
        _say(f"case: config3 cosmology {c} ({kg[c].size} k) vs complete oracle")
        o = pipeline.run(m64, plans[c])
        assert np.all(o["status"] == 0)
        e = _errors(_lensed(outs[c]), o["lensed"])
        _say(f"FP32-LOS lensed vs FP64 oracle: TT {e[0]:.1e} TE {e[1]:.1e} EE {e[2]:.1e} BB {e[3]:.1e}")
        assert max(e[:3]) < 1e-4 and e[3] < 1e-3


def test_config5_batched_gradients():
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096)
    thetas = LK.latin_hypercube(256, seed=0)
    kg = LK.uneven_k_grids(256, seed=0)
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=256, seed=1)
    _say("case: config5, 256 cosmologies, tangent gradients in chunks of 8, 16, 32, 64")
    L1, g1 = LK.tangent_gradients(m, like, thetas, k_grids=kg, chunk=64)
    assert np.all(np.isfinite(L1)) and np.all(np.isfinite(g1))
    assert np.all(np.abs(g1).max(axis=1) > 0)
    _say("case: config5, 256 cosmologies, one ODE launch")
    L2, g2 = LK.tangent_gradients(m, like, thetas, k_grids=kg, chunk=256 * 8)
    assert np.max(np.abs(L1 / L2 - 1)) < 1e-12
    scale = np.linalg.norm(g2 * thetas, axis=1, keepdims=True) / np.abs(thetas)
    assert np.max(np.abs(g1 - g2) / scale) < 1e-10


def test_power_spectra_at_redshift():
    """abcmb_pk (Pk_lin / Pk_cb at any z, abcmb/spectrum.py:268-347) against the
    oracle on the GPU's table, and P_cb / P_m -> (omega_cb / omega_m)^2 on large
    scales with a massive neutrino (the reference normalises delta_cb by omega_m)."""
    from abcmb_amd.model import Model, MassiveNeutrino
    from oracle import oracle as O
    m = Model(user_species=(MassiveNeutrino,), l_max=1000, k_axis_perturbations=np.geomspace(1e-5, 1.0, 300))
    p = m.derived(dict(P0, N_nu_massive=1, m_nu_massive=0.3))
    plan = m.prepare(p)
    _say("case: P(k, z) and P_cb(k, z) on the GPU vs the oracle")
    g = m.execute([plan])[0]
    pt = g.PT.table.cpu().numpy()
    kq = np.geomspace(1e-4, 1.5, 120)
    assert np.array_equal(m.Pk_lin(m.k_pk, 0.0, g.PT, p).cpu().numpy(), g.Pk.cpu().numpy())
    for z in (0.0, 0.5, 3.0):
        for kind, fn in ((0, m.Pk_lin), (1, m.Pk_cb)):
            got = fn(kq, z, g.PT, p).cpu().numpy()
            ref = O.pk_z(pt, plan.k, m.n_save, plan.lna_save, kq, p["A_s"], p["n_s"], m.specs["k_pivot"], z, kind,
                         p["omega_b"], p["omega_cdm"], p["omega_m"])
            assert np.max(np.abs(got / ref - 1)) < 1e-12, (z, kind)
    ratio = (m.Pk_cb(kq, 0.0, g.PT, p) / m.Pk_lin(kq, 0.0, g.PT, p)).cpu().numpy()
    w = ((p["omega_b"] + p["omega_cdm"]) / p["omega_m"]) ** 2
    assert abs(ratio[0] / w - 1) < 2e-3
    assert ratio[-1] > ratio[0]          # neutrino free streaming suppresses P_m below P_cb on small scales
