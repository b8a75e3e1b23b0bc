"""GPU discrete adjoint of the perturbation ODE (abcmb_pert_adjoint) through the C ABI.

The adjoint is pinned to the tangent-linear kernel, which the oracle pins
(test_gpu_tangent.py): for any table cotangent lam and any tangent seed s of
the host inputs (cosmology fields and the tau, x_e, T_m tables),

    <lam, J s> = <J^T lam, s>,

J the derivative of the perturbation tables with respect to those inputs at
the launch's step sequence.  The tangent kernel gives the left side along
five seed directions, the adjoint the right side for all of them at once.
Both solve their stage equations to 1e-6 (relative, per stage), so the two
sides agree to about that fraction of sum |lam * J s|; the bound is 2e-5.
k-modes whose tangent-kernel step sequence differs from the primal's (a
rounding-level accept/reject flip between the two kernels) get lam = 0.
"""
import dataclasses
from ctypes import c_void_p as C_void_p

import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu

# the AbcmbCosmo fields the perturbation equations read (bg_eval, fill_scalars,
# ic_row, all_fields)
ODE_FIELDS = ("omega_b", "omega_cdm", "omega_Lambda", "TCMB0", "YHe", "N_ur", "T_ur", "N_ncdm", "T_ncdm",
              "m_ncdm", "xe_last", "tm_last", "om", "R_nu", "R_b")


def _say(msg):
    print(msg, flush=True)


def _random_seeds(plan, rng):
    """Five smooth random tangent directions of the host inputs."""
    from abcmb_amd import abi
    from abcmb_amd.model import TangentSeeds, cosmo_doubles
    nd = abi.N_TANGENT_DIRS
    names = [f[0] for f in abi.AbcmbCosmo._fields_][:abi.COSMO_DOUBLES]
    vals = cosmo_doubles(plan.bg.cosmo)
    par = np.zeros((nd, abi.COSMO_DOUBLES))
    for f in ODE_FIELDS:
        i = names.index(f)
        par[:, i] = rng.normal(size=nd) * 1e-2 * (abs(vals[i]) + 1e-30)

    def smooth(tab):
        x = np.linspace(0.0, 1.0, tab.size)
        c = rng.normal(size=(nd, 3)) * 1e-2
        return tab[None] * (c[:, :1] + c[:, 1:2] * x[None] + c[:, 2:3] * np.sin(7.0 * x)[None])

    return TangentSeeds(par, smooth(plan.bg.tau_tab), smooth(plan.bg.xe_tab), smooth(plan.bg.tm_tab))


def _dot_identity(m, plan, seeds, rng, label):
    import torch
    outs, dpts = m.execute_tangent([plan], [seeds])
    dpt = dpts[0]
    pt = outs[0].PT.table
    prim = m.execute([plan])[0]
    same = torch.as_tensor(outs[0].steps == prim.steps, device=pt.device)
    # a table cotangent with every field weighted by its size
    rms = pt.pow(2).mean(dim=(1, 2)).sqrt()
    w = torch.where(rms > 0, 1.0 / rms, torch.zeros_like(rms))
    lam = torch.randn(pt.shape, generator=torch.Generator(device="cpu").manual_seed(int(rng.integers(1 << 30))),
                      dtype=torch.float64).to(pt.device) * w[:, None, None]
    lam = lam * same[None, None, :].to(lam.dtype)
    _say(f"case: {label}, {int(same.sum())} of {same.numel()} k-modes with the primal's step sequence")
    cot = m.ode_adjoint([plan], [lam])[0]
    assert np.array_equal(cot["steps"], prim.steps), "the adjoint's forward pass left the primal's step sequence"
    dev = pt.device
    sp = torch.as_tensor(seeds.par, device=dev)
    st = torch.as_tensor(seeds.tau, device=dev)
    sx = torch.as_tensor(seeds.xe, device=dev)
    sm = torch.as_tensor(seeds.tm, device=dev)
    for d in range(sp.shape[0]):
        lhs = float(torch.sum(lam * dpt[d]))
        rhs = float(cot["par"] @ sp[d] + cot["tau"] @ st[d] + cot["xe"] @ sx[d] + cot["tm"] @ sm[d])
        scale = float(torch.sum(torch.abs(lam * dpt[d])))
        _say(f"  direction {d}: <lam, J s> = {lhs:.12e}, <J^T lam, s> = {rhs:.12e}, "
             f"difference / sum|lam J s| = {abs(lhs - rhs) / scale:.2e}")
        assert abs(lhs - rhs) <= 2e-5 * scale


@pytest.mark.parametrize("massive", [False, True])
def test_adjoint_is_the_transpose_of_the_tangent(massive):
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.3, 12)
    m = Model(user_species=(MassiveNeutrino,) if massive else (), l_max=500, k_axis_perturbations=kk)
    plan = m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))
    rng = np.random.default_rng(11 + massive)
    _dot_identity(m, plan, _random_seeds(plan, rng), rng, f"random input directions, massive={massive}")


def test_adjoint_parameter_directions():
    """The same identity along the five parameter directions the gradients use
    (host seeds by central differences of the producer)."""
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    kk = np.geomspace(1e-4, 0.3, 12)
    m = Model(user_species=(MassiveNeutrino,), l_max=500, k_axis_perturbations=kk)
    th = LK.FIDUCIAL.copy()
    base = m.prepare(m.derived(LK.theta_to_params(th)))
    steps = LK.FD_REL_STEP * np.abs(th)
    plus, minus = [], []
    for p in LK.ODE_DIRS:
        for sgn, lst in ((1, plus), (-1, minus)):
            t = th.copy()
            t[p] += sgn * steps[p]
            lst.append(m.prepare(m.derived(LK.theta_to_params(t)), freeze=base))
    seeds = tangent_seeds(plus, minus, [steps[p] for p in LK.ODE_DIRS])
    _dot_identity(m, base, seeds, np.random.default_rng(5), "parameter directions, massive nu")


def test_lensing_adjoint_matches_finite_differences():
    """<lensed_bar, d lensed> along random directions of the unlensed spectra and
    of C_l^phiphi: central differences of abcmb_lens against the adjoint's
    <cl_bar, d cl> and <clpp_bar, d clpp>.  The lensed spectra are linear in
    the unlensed ones, so that difference is exact at any step (h = 0.1: only
    rounding is left, bound 1e-8); in C_l^phiphi they are smooth (h = 4e-3:
    truncation and rounding a few 1e-6, bound 1e-4).  The cotangent of each
    lensed spectrum is weighted by one over its largest value, so every
    multipole counts like an absolute error (weights of one over each
    multipole's value would amplify the rounding of the small high-l values)."""
    import torch
    from abcmb_amd import abi
    from abcmb_amd.model import Model
    m = Model(lensing=True, l_max=1000, k_axis_perturbations=np.geomspace(1e-4, 0.3, 8))
    plan = m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))
    out = m.execute([plan])[0]
    lib = abi.hip_lib()
    dev = torch.device("cuda")
    nl = m.lensing_ells.size
    D = m._device_constants()
    ds = m._downstream([plan], [out.PT.table], [D["k_pt"]])
    cl, clpp = ds["cl_full"].contiguous(), ds["clpp"].contiguous()
    ws = torch.empty(lib.abcmb_lens_workspace_size(1, m.mu.size, nl) // 8 + 1, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream

    def lensed(c, p):
        o = torch.empty((1, 4, nl), dtype=torch.float64, device=dev)
        abi.check(lib.abcmb_lens(1, c.contiguous().data_ptr(), p.contiguous().data_ptr(), nl, D["mu"].data_ptr(),
                                 D["w"].data_ptr(), m.mu.size, o.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream),
                  "abcmb_lens")
        return o

    base = lensed(cl, clpp)
    g = torch.Generator(device="cpu").manual_seed(3)
    dcl = torch.randn(cl.shape, generator=g, dtype=torch.float64).to(dev) * cl
    dpp = torch.randn(clpp.shape, generator=g, dtype=torch.float64).to(dev) * clpp
    lb = (torch.randn((1, 4, nl), generator=g, dtype=torch.float64).to(dev) /
          base.abs().amax(dim=2, keepdim=True)).contiguous()
    clb = torch.empty_like(cl)
    ppb = torch.empty_like(clpp)
    wsa = torch.empty(lib.abcmb_lens_adjoint_workspace_size(1, m.mu.size, nl) // 8 + 1, dtype=torch.float64,
                      device=dev)
    abi.check(lib.abcmb_lens_adjoint(1, cl.data_ptr(), clpp.data_ptr(), nl, D["mu"].data_ptr(), D["w"].data_ptr(),
                                     m.mu.size, lb.data_ptr(), clb.data_ptr(), ppb.data_ptr(), wsa.data_ptr(),
                                     wsa.numel() * 8, stream), "abcmb_lens_adjoint")
    for name, h, bound, dc, dp, ad in (("unlensed C_l", 0.1, 1e-8, dcl, 0 * dpp, float(torch.sum(clb * dcl))),
                                       ("C_l^phiphi", 4e-3, 1e-4, 0 * dcl, dpp, float(torch.sum(ppb * dpp)))):
        fd = float(torch.sum(lb * (lensed(cl + h * dc, clpp + h * dp) - lensed(cl - h * dc, clpp - h * dp)))) / (2 * h)
        _say(f"case: lensing adjoint along {name}: differences {fd:.12e}, adjoint {ad:.12e}, "
             f"relative {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= bound * abs(ad)


@pytest.mark.parametrize("l_max", [1000, 2500])
def test_adjoint_gradients_match_tangent_gradients(l_max):
    """The reverse-mode likelihood gradient (downstream adjoint + ODE adjoint,
    contracted with the host seeds) against the tangent-linear path
    (``tangent_gradients``, pinned to the oracle in test_gpu_tangent.py), on
    two cosmologies with massive neutrinos: the two linearise the same
    discretisation, apart from the tangent path's central differences of the
    lensing (relative step 1e-4) and the two ODE linearisations' stage-solve
    tolerance (1e-6); bound 1e-4 of the gradient norm."""
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.5, 24)
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=l_max, k_axis_perturbations=kk)
    th = LK.latin_hypercube(2, seed=4)
    # the mock at the fiducial point, away from both: at a mock's own parameters
    # the gradient is a sum of random-signed terms that nearly cancels, and
    # both linearisations' 1e-6 stage-solve tolerances show up magnified there
    base = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [base.ClTT.cpu().numpy(), base.ClTE.cpu().numpy(), base.ClEE.cpu().numpy()],
                                 n_mocks=2, seed=1)
    _say(f"case: adjoint vs tangent gradients, l_max={l_max}")
    La, ga = LK.adjoint_gradients(m, like, th)
    Lt, gt = LK.tangent_gradients(m, like, th)
    _say(f"L adjoint {La}, L tangent {Lt}")
    # L from the primal kernel's tables and from the tangent kernel's values, which
    # differ by up to the Newton tolerance (5e-6 of a field) where a chord test flips
    assert np.allclose(La, Lt, rtol=1e-5, atol=0)
    scale = np.linalg.norm(gt * th, axis=1, keepdims=True) / np.abs(th)
    err = np.abs(ga - gt) / scale
    _say(f"gradient difference / norm:\n{np.array2string(err, precision=2)}")
    _say(f"adjoint:\n{np.array2string(ga, precision=6)}\ntangent:\n{np.array2string(gt, precision=6)}")
    assert np.all(err <= 1e-4)


def _smooth_table_direction(pt, seed):
    """A smooth random direction of a perturbation table: each field times a
    random smooth function of (ln a index, k index)."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    nf, ns, nk = pt.shape
    a = torch.linspace(0, 1, ns, dtype=torch.float64)[:, None]
    b = torch.linspace(0, 1, nk, dtype=torch.float64)[None, :]
    c = torch.randn((nf, 4), generator=g, dtype=torch.float64)
    shape = (c[:, 0, None, None] + c[:, 1, None, None] * a + c[:, 2, None, None] * b +
             c[:, 3, None, None] * torch.sin(5 * a + 3 * b))
    return (pt * shape.to(pt.device) * 1e-2).contiguous()


@pytest.mark.parametrize("l_max", [1000, 2500])
def test_los_adjoint_is_the_transpose_of_the_tangent_los(l_max):
    """<cl_bar, dC> = <pt_bar, dT> for the sources + LOS + k-quadrature, with the
    tangent-linear LOS (abcmb_los_tangent) fed the sources of a table direction
    dT (the sources are linear in the table: dsrc = sources(dT) exactly at a
    fixed background, dchi = 0) against abcmb_los_adjoint."""
    import torch
    from abcmb_amd import abi
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.5, 24)
    m = Model(user_species=(MassiveNeutrino,), l_max=l_max, k_axis_perturbations=kk)
    plan = m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))
    out = m.execute([plan])[0]
    lib = abi.hip_lib()
    D = m._device_constants()
    dev = torch.device("cuda")
    pt = out.PT.table.contiguous()
    kd = D["k_pt"]
    nd = abi.N_TANGENT_DIRS
    dts = [_smooth_table_direction(pt, 20 + d) for d in range(nd)]
    src = m._run_sources([plan], [pt], [kd])
    dsrc = torch.stack([m._run_sources([plan], [dt_], [kd])[0] for dt_ in dts])[None].contiguous()
    n_tau, nkt, n_ell = m.n_save - 1, m.k_transfer.size, m.ell_idx.size
    dchi = torch.zeros((1, nd, n_tau), dtype=torch.float64, device=dev)
    _, _, stage = m._staging()
    inp = m._downstream_inputs([plan], [pt], [kd], stage.add)
    stage.flush()
    cos, bgv = inp["cos"], inp["bgv"]
    nv = 1 + nd + 2
    raw = torch.empty((1, nv, 3, n_ell), dtype=torch.float64, device=dev)
    ws = torch.empty(lib.abcmb_los_tangent_workspace_size(1, nd, n_ell, nkt) // 8 + 1, dtype=torch.float64,
                     device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    abi.check(lib.abcmb_los_tangent(cos.data_ptr(), 1, src.data_ptr(), dsrc.data_ptr(), dchi.data_ptr(), nd, n_tau,
                                    bgv.data_ptr(), D["kt"].data_ptr(), nkt, D["bes"].data_ptr(), D["rng"].data_ptr(),
                                    D["bes_l"].data_ptr(), m.bes_l.size, m.bes_tab.shape[2], D["ell_idx"].data_ptr(),
                                    n_ell, 0, raw.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream),
              "abcmb_los_tangent")
    g = torch.Generator(device="cpu").manual_seed(9)
    cb = (torch.randn((1, 3, n_ell), generator=g, dtype=torch.float64).to(dev) /
          raw[:, 0].abs().amax(dim=2, keepdim=True)).contiguous()
    ptb = torch.zeros_like(pt)
    ptrs = torch.as_tensor(np.array([ptb.data_ptr()], dtype=np.int64), device=dev)
    bgvb = torch.empty((1, 6, n_tau), dtype=torch.float64, device=dev)
    scb = torch.empty((1, 3), dtype=torch.float64, device=dev)
    wsa = torch.empty(lib.abcmb_los_adjoint_workspace_size(1, kk.size, m.n_save, n_ell, nkt) // 8 + 1,
                      dtype=torch.float64, device=dev)
    abi.check(lib.abcmb_los_adjoint(cos.data_ptr(), 1, kk.size, m.n_save, bgv.data_ptr(), D["kt"].data_ptr(), nkt,
                                    m.scales.ctypes.data_as(C_void_p), D["bes"].data_ptr(), D["rng"].data_ptr(),
                                    D["bes_l"].data_ptr(), m.bes_l.size, m.bes_tab.shape[2], D["ell_idx"].data_ptr(),
                                    n_ell, cb.data_ptr(), ptrs.data_ptr(), bgvb.data_ptr(), scb.data_ptr(),
                                    wsa.data_ptr(), wsa.numel() * 8, stream), "abcmb_los_adjoint")
    torch.cuda.synchronize()
    for d in range(nd):
        lhs = float(torch.sum(cb[0] * raw[0, 1 + d]))
        rhs = float(torch.sum(ptb * dts[d]))
        _say(f"case: LOS adjoint l_max={l_max}, direction {d}: <cb, dC> {lhs:.12e}, <ptb, dT> {rhs:.12e}, "
             f"relative {abs(lhs - rhs) / abs(lhs):.2e}")
        assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
    # d/d ln A_s and d/d n_s of the sampled C_l against the adjoint's A_s, n_s cotangents
    As, nS = plan.bg.cosmo.A_s, plan.bg.cosmo.n_s
    for q, col, f in ((0, 1 + nd, As), (1, 2 + nd, 1.0)):
        lhs = float(torch.sum(cb[0] * raw[0, col]))
        rhs = float(scb[0, q]) * f
        _say(f"  primordial {q}: {lhs:.12e} vs {rhs:.12e}")
        assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_limber_adjoint_matches_differences():
    """C_l^phiphi is quadratic in the delta_m field, so central differences along a
    table direction are exact up to rounding: <clpp_bar, dClpp> against
    <pt_bar, dT> from abcmb_lensing_clpp_adjoint."""
    import torch
    from abcmb_amd import abi
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.5, 24)
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, k_axis_perturbations=kk)
    plan = m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))
    out = m.execute([plan])[0]
    D = m._device_constants()
    dev = torch.device("cuda")
    pt = out.PT.table.contiguous()
    dT = _smooth_table_direction(pt, 31)
    h = 0.5
    cp = m.clpp_of([plan], [(pt + h * dT).contiguous()], [D["k_pt"]])
    cm = m.clpp_of([plan], [(pt - h * dT).contiguous()], [D["k_pt"]])
    c0 = m.clpp_of([plan], [pt], [D["k_pt"]])
    g = torch.Generator(device="cpu").manual_seed(2)
    cb = (torch.randn(c0.shape, generator=g, dtype=torch.float64).to(dev) / c0.abs().amax()).contiguous()
    lhs = float(torch.sum(cb * (cp - cm))) / (2 * h)
    from abcmb_amd import abi
    lib = abi.hip_lib()
    _, _, stage = m._staging()
    inp = m._downstream_inputs([plan], [pt], [D["k_pt"]], stage.add)
    stage.flush()
    ptb = torch.zeros_like(pt)
    ptrs = torch.as_tensor(np.array([ptb.data_ptr()], dtype=np.int64), device=dev)
    abi.check(lib.abcmb_lensing_clpp_adjoint(inp["cos"].data_ptr(), 1, inp["lens"].data_ptr(), inp["lens"].shape[2],
                                             D["lens_ells"].data_ptr(), m.lensing_ells.size, cb.data_ptr(),
                                             ptrs.data_ptr(), torch.cuda.current_stream().cuda_stream),
              "abcmb_lensing_clpp_adjoint")
    rhs = float(torch.sum(ptb * dT))
    _say(f"case: Limber adjoint: differences {lhs:.12e}, adjoint {rhs:.12e}, relative {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)


def test_los_adjoint_background_part():
    """The LOS background's part of abcmb_los_adjoint (bgv_bar, and tau0 in
    scal_bar) against the tangent-linear LOS along a direction of the
    background vectors and of tau0: the sources' derivative by central
    differences of abcmb_sources (relative step 1e-5; they are products of the
    background and the fields), the distance's dchi = dtau0 - dtau(s)."""
    import torch
    from abcmb_amd import abi
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.5, 24)
    m = Model(user_species=(MassiveNeutrino,), l_max=2500, k_axis_perturbations=kk)
    plan = m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))
    out = m.execute([plan])[0]
    lib = abi.hip_lib()
    D = m._device_constants()
    dev = torch.device("cuda")
    pt = out.PT.table.contiguous()
    kd = D["k_pt"]
    nd = abi.N_TANGENT_DIRS
    n_tau, nkt, n_ell = m.n_save - 1, m.k_transfer.size, m.ell_idx.size
    rng = np.random.default_rng(8)
    x = np.linspace(0.0, 1.0, n_tau)
    dB = [plan.bgv * (rng.normal(size=(6, 1)) * 1e-2 + rng.normal(size=(6, 1)) * 1e-2 * np.sin(4 * x)[None])
          for _ in range(nd)]
    dtau0 = rng.normal(size=nd) * 1e-2 * plan.bg.cosmo.tau0
    h = 1e-5
    dsrc, dchi = [], []
    for d in range(nd):
        pp = dataclasses.replace(plan, bgv=plan.bgv + h * dB[d])
        pm = dataclasses.replace(plan, bgv=plan.bgv - h * dB[d])
        sp_ = m._run_sources([pp], [pt], [kd])[0]
        sm_ = m._run_sources([pm], [pt], [kd])[0]
        dsrc.append((sp_ - sm_) / (2 * h))
        dchi.append(dtau0[d] - dB[d][0])
    dsrc = torch.stack(dsrc)[None].contiguous()
    dchi_t = torch.as_tensor(np.stack(dchi)[None], device=dev).contiguous()
    src = m._run_sources([plan], [pt], [kd])
    _, _, stage = m._staging()
    inp = m._downstream_inputs([plan], [pt], [kd], stage.add)
    stage.flush()
    cos, bgv = inp["cos"], inp["bgv"]
    raw = torch.empty((1, 1 + nd + 2, 3, n_ell), dtype=torch.float64, device=dev)
    ws = torch.empty(lib.abcmb_los_tangent_workspace_size(1, nd, n_ell, nkt) // 8 + 1, dtype=torch.float64,
                     device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    abi.check(lib.abcmb_los_tangent(cos.data_ptr(), 1, src.data_ptr(), dsrc.data_ptr(), dchi_t.data_ptr(), nd, n_tau,
                                    bgv.data_ptr(), D["kt"].data_ptr(), nkt, D["bes"].data_ptr(), D["rng"].data_ptr(),
                                    D["bes_l"].data_ptr(), m.bes_l.size, m.bes_tab.shape[2], D["ell_idx"].data_ptr(),
                                    n_ell, 0, raw.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream),
              "abcmb_los_tangent")
    g = torch.Generator(device="cpu").manual_seed(4)
    cb = (torch.randn((1, 3, n_ell), generator=g, dtype=torch.float64).to(dev) /
          raw[:, 0].abs().amax(dim=2, keepdim=True)).contiguous()
    ptb = torch.zeros_like(pt)
    ptrs = torch.as_tensor(np.array([ptb.data_ptr()], dtype=np.int64), device=dev)
    bgvb = torch.empty((1, 6, n_tau), dtype=torch.float64, device=dev)
    scb = torch.empty((1, 3), dtype=torch.float64, device=dev)
    wsa = torch.empty(lib.abcmb_los_adjoint_workspace_size(1, kk.size, m.n_save, n_ell, nkt) // 8 + 1,
                      dtype=torch.float64, device=dev)
    abi.check(lib.abcmb_los_adjoint(cos.data_ptr(), 1, kk.size, m.n_save, bgv.data_ptr(), D["kt"].data_ptr(), nkt,
                                    m.scales.ctypes.data_as(C_void_p), D["bes"].data_ptr(), D["rng"].data_ptr(),
                                    D["bes_l"].data_ptr(), m.bes_l.size, m.bes_tab.shape[2], D["ell_idx"].data_ptr(),
                                    n_ell, cb.data_ptr(), ptrs.data_ptr(), bgvb.data_ptr(), scb.data_ptr(),
                                    wsa.data_ptr(), wsa.numel() * 8, stream), "abcmb_los_adjoint")
    bb = bgvb[0].cpu().numpy()
    for d in range(nd):
        lhs = float(torch.sum(cb[0] * raw[0, 1 + d]))
        rhs = float(np.sum(bb * dB[d]) + float(scb[0, 2]) * dtau0[d])
        _say(f"case: LOS adjoint background direction {d}: tangent {lhs:.12e}, adjoint {rhs:.12e}, "
             f"relative {abs(lhs - rhs) / abs(lhs):.2e}")
        assert abs(lhs - rhs) <= 1e-7 * abs(lhs)
