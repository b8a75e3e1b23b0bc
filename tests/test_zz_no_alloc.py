"""abcmb_pert_fwd_noalloc (the allocation-free perturbation launch) gives what
abcmb_pert_fwd with a caller-owned workspace gives, bit for bit, on two calls
in a row (the second smaller, so the fixed scratch holds the first call's
contents when it starts), and refuses a launch beyond its fixed scratch."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _launch(m, lib, plans, k_sub, noalloc):
    import torch
    from abcmb_amd import abi
    _, dev, stage = m._staging()
    dt = stage.add
    ns, nk = m.n_save, m.k_pt.size
    kdev = [dt(m.k_pt) for _ in plans]
    pts = [torch.full((abi.N_PT_FIELDS, ns, nk), np.nan, dtype=torch.float64, device=dev) for _ in plans]
    cos_dev = m._struct_array(torch, dev, m._cosmo_array(plans, pts, kdev, dt), dt)
    pairs = [(ci, ik) for ci in range(len(plans)) for ik in k_sub]
    items = [abi.AbcmbPertItem(k=m.k_pt[ik], lna_start=plans[ci].lna_start[ik], rtol=plans[ci].rtol[ik],
                               atol=plans[ci].atol[ik], cosmo=ci, ik=ik) for ci, ik in pairs]
    items_dev = m._struct_array(torch, dev, items, dt)
    n = len(items)
    status = torch.full((n,), -1, dtype=torch.int32, device=dev)
    steps = torch.zeros(n, dtype=torch.int32, device=dev)
    iters = torch.zeros(n, dtype=torch.int32, device=dev)
    stage.flush()
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = lib.abcmb_pert_fwd_noalloc(cos_dev.data_ptr(), len(plans), items_dev.data_ptr(), n, C.byref(m.settings),
                                        ns, status.data_ptr(), steps.data_ptr(), iters.data_ptr(), None, stream)
        abi.check(rc, "abcmb_pert_fwd_noalloc")
    else:
        ws = torch.empty(lib.abcmb_pert_workspace_size(len(plans), ns, 0) // 8 + 1, dtype=torch.float64, device=dev)
        rc = lib.abcmb_pert_fwd(cos_dev.data_ptr(), len(plans), items_dev.data_ptr(), n, C.byref(m.settings), ns,
                                status.data_ptr(), steps.data_ptr(), iters.data_ptr(), None, ws.data_ptr(),
                                ws.numel() * 8, stream)
        abi.check(rc, "abcmb_pert_fwd")
    torch.cuda.synchronize()
    return pts, status, steps, iters


def test_noalloc_launch_matches_workspace_launch():
    import torch
    from abcmb_amd import abi
    from abcmb_amd import likelihood as LK
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=False, l_max=500,
              k_axis_perturbations=np.geomspace(1e-4, 0.3, 12))
    lib = abi.hip_lib()
    abi.check(lib.abcmb_device_init(), "abcmb_device_init")
    thetas = LK.latin_hypercube(3, seed=5)
    plans = m.prepare_many([m.derived(LK.theta_to_params(t)) for t in thetas])
    cases = [(plans, list(range(12))), (plans[2:], [0, 3, 6, 9, 11])]   # the second call is smaller
    for i, (pl, ks) in enumerate(cases):
        _say(f"case: no-alloc call {i}: {len(pl)} cosmologies x {len(ks)} k-modes (massive nu, n=100)")
        a = _launch(m, lib, pl, ks, noalloc=False)
        b = _launch(m, lib, pl, ks, noalloc=True)
        assert int(a[1].abs().max()) == 0 and int(b[1].abs().max()) == 0
        for x, y in zip(a[1:], b[1:]):
            assert torch.equal(x, y)
        for pa, pb in zip(a[0], b[0]):
            assert torch.equal(pa[:, :, ks], pb[:, :, ks])
            assert torch.isnan(pb[:, :, [k for k in range(12) if k not in ks]]).all()


def test_noalloc_refuses_launch_beyond_its_scratch():
    from abcmb_amd import abi
    from abcmb_amd.model import Model
    m = Model(lensing=False, l_max=300, k_axis_perturbations=np.geomspace(1e-4, 0.1, 4))
    lib = abi.hip_lib()
    n_cosmo, ns = 200, 500
    assert lib.abcmb_pert_workspace_size(n_cosmo, ns, 0) > 8 << 20
    one = abi.AbcmbPertItem()
    # rejected on the sizes alone, before any pointer is read or any GPU work is enqueued
    rc = lib.abcmb_pert_fwd_noalloc(C.addressof(one), n_cosmo, C.addressof(one), 1, C.byref(m.settings), ns,
                                    C.addressof(one), C.addressof(one), None, None, None)
    assert rc == 1
    assert b"fixed scratch" in lib.abcmb_last_error()
