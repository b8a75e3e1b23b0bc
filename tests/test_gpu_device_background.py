"""Background completion on the GPU (``Model(device_background=True)``):

  - abcmb_los_background (tau, g, g', aH, exp(-kappa), aH' aH with the optical
    depth integrated on the device) equals the host producer's LOS vectors to
    1e-13 of each vector's maximum for three Latin-hypercube cosmologies
    (measured 2e-15: the panel sums add in another order), and its tangents
    equal central differences of the host vectors (relative step 1e-4, grids
    frozen) to 2e-7 of each maximum (measured 2e-8);
  - a model that leaves reionization and the LOS background to the GPU gives
    the host-produced model's x_e tables (1e-12), z_reion (1e-10), LOS vectors
    (1e-13) and lensed C_l (1e-10; measured 5e-12) for the same cosmologies;
  - its tangent gradients, with the x_e and z_reion seeds from the
    implicit-function derivatives of the reionization root instead of host
    finite differences, equal the host-seeded gradients to 1e-8 of the norm
    (measured 3e-11).
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_los_background_on_device_matches_host():
    import torch
    from abcmb_amd import abi
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    m = Model(user_species=(MassiveNeutrino,), l_max=1000)
    lib = abi.hip_lib()
    dev = torch.device("cuda")
    d = lambda a, dt=torch.float64: torch.as_tensor(np.array(a), dtype=dt, device=dev)
    thetas = LK.latin_hypercube(3, seed=52)
    plans = [m.prepare(m.derived(LK.theta_to_params(t))) for t in thetas]
    nd, nl = abi.N_TANGENT_DIRS, m.n_save - 1
    cos, tans, keep, fd = [], [], [], []
    for th, pl in zip(thetas, plans):
        steps = LK.FD_REL_STEP * np.abs(th)
        plus, minus = [], []
        for p in LK.ODE_DIRS:
            for sgn, lst in ((1, plus), (-1, minus)):
                t = th.copy()
                t[p] += sgn * steps[p]
                lst.append(m.prepare(m.derived(LK.theta_to_params(t)), freeze=pl))
        h = [steps[p] for p in LK.ODE_DIRS]
        sd = tangent_seeds(plus, minus, h)
        fd.append(np.stack([(a.bgv - b.bgv) / (2 * s) for a, b, s in zip(plus, minus, h)]))
        c = abi.AbcmbCosmo.from_buffer_copy(pl.bg.cosmo)
        tt = [d(pl.bg.tau_tab), d(pl.bg.xe_tab), d(pl.bg.tm_tab), d(sd.par), d(sd.tau), d(sd.xe), d(sd.tm)]
        keep += tt
        c.tau_tab, c.xe_tab, c.tm_tab = (x.data_ptr() for x in tt[:3])
        cos.append(c)
        tans.append(abi.AbcmbTangent(par=tt[3].data_ptr(), tau_tab=tt[4].data_ptr(), xe_tab=tt[5].data_ptr(),
                                     tm_tab=tt[6].data_ptr(), pt=None))
    n = len(plans)
    cos_d = d(np.frombuffer(bytes((abi.AbcmbCosmo * n)(*cos)), np.uint8), torch.uint8)
    tan_d = d(np.frombuffer(bytes((abi.AbcmbTangent * n)(*tans)), np.uint8), torch.uint8)
    lna = d(np.stack([pl.lna_save[:-1] for pl in plans]))
    out = torch.empty((n, 6, nl), dtype=torch.float64, device=dev)
    dout = torch.empty((n, nd, 6, nl), dtype=torch.float64, device=dev)
    out0 = torch.empty((n, 6, nl), dtype=torch.float64, device=dev)
    _say("case: LOS background of 3 cosmologies on the GPU, with and without tangents")
    stream = torch.cuda.current_stream().cuda_stream
    abi.check(lib.abcmb_los_background(cos_d.data_ptr(), tan_d.data_ptr(), n, lna.data_ptr(), nl, out.data_ptr(),
                                       dout.data_ptr(), stream), "abcmb_los_background")
    abi.check(lib.abcmb_los_background(cos_d.data_ptr(), None, n, lna.data_ptr(), nl, out0.data_ptr(), None,
                                       stream), "abcmb_los_background")
    torch.cuda.synchronize()
    o, o0, do = out.cpu().numpy(), out0.cpu().numpy(), dout.cpu().numpy()
    assert np.array_equal(o, o0)
    for c, pl in enumerate(plans):
        err = np.max(np.abs(o[c] - pl.bgv), axis=1) / np.max(np.abs(pl.bgv), axis=1)
        derr = np.max(np.abs(do[c] - fd[c]), axis=2) / (np.max(np.abs(fd[c]), axis=2) + 1e-300)
        _say(f"cosmology {c}: values {np.array2string(err, precision=1)}; tangents worst {derr.max():.1e}")
        assert err.max() < 1e-13
        # aH and tau do not move with tau_reion: compare there on the other directions' scale
        live = np.max(np.abs(fd[c]), axis=2) > 1e-12 * np.max(np.abs(fd[c]), axis=(0, 2))[None, :]
        assert np.max(derr[live]) < 2e-7


def _cl(o):
    return [x.cpu().numpy() for x in (o.ClTT, o.ClTE, o.ClEE, o.ClBB)]


def test_device_background_model_matches_host_model():
    from abcmb_amd.model import Model, MassiveNeutrino
    kw = dict(user_species=(MassiveNeutrino,), lensing=True, l_max=1000)
    mh, md = Model(**kw), Model(device_background=True, **kw)
    thetas = LK.latin_hypercube(3, seed=53)
    ph = mh.prepare_many([mh.derived(LK.theta_to_params(t)) for t in thetas])
    pd = md.prepare_many([md.derived(LK.theta_to_params(t)) for t in thetas])
    assert all(p.bgv is None and p.bg.z_reion == 0.0 for p in pd)
    _say("case: 3 cosmologies, reionization and LOS background on the GPU vs on the host")
    oh, od = mh.execute(ph), md.execute(pd)
    for a, b, x, y in zip(ph, pd, oh, od):
        assert abs(b.bg.z_reion / a.bg.z_reion - 1) < 1e-10
        assert np.max(np.abs(b.bg.xe_tab - a.bg.xe_tab)) < 1e-12
        assert np.max(np.abs(b.bgv - a.bgv) / np.max(np.abs(a.bgv), axis=1, keepdims=True)) < 1e-13
        ch, cd = _cl(x), _cl(y)
        e = [np.max(np.abs(cd[i] - ch[i]) / (np.max(np.abs(ch[i])) if i == 1 else np.abs(ch[i]))) for i in range(4)]
        _say(f"lensed C_l, device vs host background: TT {e[0]:.1e} TE {e[1]:.1e} EE {e[2]:.1e} BB {e[3]:.1e}")
        assert max(e) < 1e-10


def test_device_background_gradients_match_host_seeds():
    from abcmb_amd.model import Model, MassiveNeutrino
    kw = dict(user_species=(MassiveNeutrino,), lensing=True, l_max=1000)
    mh, md = Model(**kw), Model(device_background=True, **kw)
    thetas = LK.latin_hypercube(3, seed=54)
    kg = [k[::4] for k in LK.uneven_k_grids(3, seed=54, hi=0.3)]
    fid = mh.execute([mh.prepare(mh.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    obs = [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()]
    lh = LK.LensedClLikelihood(mh, obs, n_mocks=3, seed=1)
    ld = LK.LensedClLikelihood(md, obs, n_mocks=3, seed=1)
    _say("case: tangent gradients of 3 cosmologies, host FD seeds vs device implicit-function seeds")
    Lh, gh = LK.tangent_gradients(mh, lh, thetas, k_grids=kg)
    Ld, gd = LK.tangent_gradients(md, ld, thetas, k_grids=kg)
    scale = np.linalg.norm(gh * thetas, axis=1, keepdims=True) / np.abs(thetas)
    err = np.abs(gd - gh) / scale
    _say(f"L {np.max(np.abs(Ld / Lh - 1)):.1e}; gradient per parameter (of the norm) "
         f"{np.array2string(err.max(axis=0), precision=1)}")
    assert np.max(np.abs(Ld / Lh - 1)) < 1e-8
    assert err.max() < 1e-8
