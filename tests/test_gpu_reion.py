"""Reionization on the GPU (abcmb_reionize) against the host producer:
  - on recombination-only x_e tables, the device root of tau_reion(z_reion)
    and the reionized table equal the host producer's (z to 1e-10, x_e to
    1e-12), for three Latin-hypercube cosmologies;
  - the implicit-function derivatives of z_reion and the tangents of the
    reionized table along the five ODE directions equal central differences
    of the host producer (relative step 1e-4, grids frozen) to 1e-6.
"""
import ctypes as C

import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


class _Plan:
    def __init__(self, bg):
        self.bg = bg


def test_reionization_on_device_matches_host():
    import torch
    from abcmb_amd import abi
    from abcmb_amd.background import Background
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    m = Model(user_species=(MassiveNeutrino,), lensing=False, l_max=1000)
    lib = abi.hip_lib()
    hl = abi.host_lib()
    n_rec = hl.abcmb_host_recomb_points()
    thetas = LK.latin_hypercube(3, seed=51)
    dev = torch.device("cuda")
    d = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)

    def bg(th, full, base=None):
        p = m.derived(LK.theta_to_params(th))
        kw = {}
        if base is not None:
            kw = dict(rec_switch=base.rec_switch, lna_rec=base.lna_rec, lna_transfer_start=base.lna_transfer_start)
        return Background(p, True, reionization=full, **kw)

    full = [bg(t, True) for t in thetas]
    rec = [bg(t, False, b) for t, b in zip(thetas, full)]
    keep = []
    cos, tans, dtarget, fd_z, fd_xe = [], [], [], [], []
    for c, (th, bf, br) in enumerate(zip(thetas, full, rec)):
        steps = LK.FD_REL_STEP * np.abs(th)
        plus_r, minus_r, plus_f, minus_f, sd = [], [], [], [], []
        for p in LK.ODE_DIRS:
            for sgn, lr, lf in ((1, plus_r, plus_f), (-1, minus_r, minus_f)):
                t = th.copy()
                t[p] += sgn * steps[p]
                lr.append(_Plan(bg(t, False, bf)))
                lf.append(bg(t, True, bf))
            sd.append(steps[p])
        seeds = tangent_seeds(plus_r, minus_r, sd)
        fd_z.append([(pp.z_reion - pm.z_reion) / (2 * h) for pp, pm, h in zip(plus_f, minus_f, sd)])
        fd_xe.append([(pp.xe_tab - pm.xe_tab) / (2 * h) for pp, pm, h in zip(plus_f, minus_f, sd)])
        dtarget.append([1.0 if LK.NAMES[p] == "tau_reion" else 0.0 for p in LK.ODE_DIRS])
        cc = abi.AbcmbCosmo.from_buffer_copy(br.cosmo)
        t_tau, t_xe, t_tm = d(br.tau_tab), d(br.xe_tab), d(br.tm_tab)
        cc.tau_tab, cc.xe_tab, cc.tm_tab = t_tau.data_ptr(), t_xe.data_ptr(), t_tm.data_ptr()
        sp, stau, sxe, stm = d(seeds.par), d(seeds.tau), d(seeds.xe), d(seeds.tm)
        keep += [t_tau, t_xe, t_tm, sp, stau, sxe, stm]
        cos.append(cc)
        tans.append(abi.AbcmbTangent(par=sp.data_ptr(), tau_tab=stau.data_ptr(), xe_tab=sxe.data_ptr(),
                                     tm_tab=stm.data_ptr(), pt=None))
    n = len(cos)
    xe_n = rec[0].xe_tab.size
    cos_d = d(np.frombuffer(bytes((abi.AbcmbCosmo * n)(*cos)), np.uint8), torch.uint8)
    tan_d = d(np.frombuffer(bytes((abi.AbcmbTangent * n)(*tans)), np.uint8), torch.uint8)
    reion_d = d(np.frombuffer(bytes((abi.AbcmbReion * n)(*[b.reion for b in rec])), np.uint8), torch.uint8)
    dtar = d(np.array(dtarget))
    nd = abi.N_TANGENT_DIRS
    xe_out = torch.zeros((n, xe_n), dtype=torch.float64, device=dev)
    dxe_out = torch.zeros((n, nd, xe_n), dtype=torch.float64, device=dev)
    z_out = torch.zeros(n, dtype=torch.float64, device=dev)
    dz_out = torch.zeros((n, nd), dtype=torch.float64, device=dev)
    st = torch.full((n,), -1, dtype=torch.int32, device=dev)
    _say("case: reionization of 3 cosmologies on the GPU, with tangents")
    abi.check(lib.abcmb_reionize(cos_d.data_ptr(), tan_d.data_ptr(), n, reion_d.data_ptr(), dtar.data_ptr(), n_rec,
                                 xe_out.data_ptr(), dxe_out.data_ptr(), z_out.data_ptr(), dz_out.data_ptr(),
                                 st.data_ptr(), torch.cuda.current_stream().cuda_stream), "abcmb_reionize")
    torch.cuda.synchronize()
    assert int(st.abs().max()) == 0
    z, xe, dz, dxe = (x.cpu().numpy() for x in (z_out, xe_out, dz_out, dxe_out))
    for c in range(n):
        assert abs(z[c] / full[c].z_reion - 1) < 1e-10
        assert np.max(np.abs(xe[c] - full[c].xe_tab)) < 1e-12
        ref_dz = np.array(fd_z[c])
        _say(f"dz_reion/dtheta device {np.array2string(dz[c], precision=4)} host FD "
             f"{np.array2string(ref_dz, precision=4)}")
        assert np.max(np.abs(dz[c] - ref_dz)) < 1e-6 * np.max(np.abs(ref_dz))
        ref_dxe = np.array(fd_xe[c])
        assert np.max(np.abs(dxe[c] - ref_dxe)) < 1e-6 * np.max(np.abs(ref_dxe))
