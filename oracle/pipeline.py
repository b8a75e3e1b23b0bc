"""End-to-end oracle run of one prepared cosmology -- TEST INFRASTRUCTURE ONLY.

Takes the host-side ``CosmoPlan`` that ``abcmb_amd.Model.prepare`` builds
(the same background tables the GPU path reads) and evaluates the whole
perturbation + LOS + lensing path with the CPU oracle, so the GPU result can
be compared with it piece by piece.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from abcmb_amd import abi
from . import oracle as O


def run(model, plan, want_states=False, lensing=None, k_subset=None):
    """Returns dict(pt, status, steps, states, cl_raw, cl_full, clpp, lensed, pk)."""
    kk = np.ascontiguousarray(plan.k if plan.k is not None else model.k_pt)
    nk, ns = kk.size, model.n_save
    pt = np.zeros((abi.N_PT_FIELDS, ns, nk))
    c = abi.AbcmbCosmo.from_buffer_copy(plan.bg.cosmo)
    lna = np.ascontiguousarray(plan.lna_save)
    c.lna_save = lna.ctypes.data
    c.k = kk.ctypes.data
    c.pt = pt.ctypes.data
    c.n_save, c.nk = ns, nk
    cos = (abi.AbcmbCosmo * 1)(c)
    ks = np.arange(nk) if k_subset is None else np.asarray(k_subset)
    items = (abi.AbcmbPertItem * ks.size)()
    for m, ik in enumerate(ks):
        items[m] = abi.AbcmbPertItem(k=kk[ik], lna_start=plan.lna_start[ik], rtol=plan.rtol[ik],
                                     atol=plan.atol[ik], cosmo=0, ik=int(ik))
    status, steps, states = O.pert(cos, items, model.settings, want_states=want_states)
    out = dict(pt=pt, status=status, steps=steps, states=states, cosmo=c, keep=(lna, kk))
    if k_subset is not None:
        return out
    p = plan.params
    src = O.sources(pt, kk, ns, plan.bgv, plan.lna_save, model.k_transfer, model.scales)
    cl_raw = O.los_cl(src, plan.bgv[0], plan.bg.tau0, model.k_transfer, model.bes_tab, model.bes_rng, model.bes_l,
                      model.ell_idx, p["A_s"], p["n_s"], model.specs["k_pivot"])
    cl_full = cl_raw @ model.ell_op.T
    out.update(src=src, cl_raw=cl_raw, cl_full=cl_full)
    do_lens = model.specs["lensing"] if lensing is None else lensing
    if do_lens:
        clpp = O.clpp(c, pt, plan.lens_lna, plan.lens_aH, plan.lens_tau, plan.chi_rec,
                      model.lensing_ells.astype(np.float64))
        lensed = O.lens(cl_full, clpp, model.mu, model.w)
        out.update(clpp=clpp, lensed=lensed)
    out["pk"] = O.pk(pt, kk, ns, plan.lna_save, model.k_pk, p["A_s"], p["n_s"], model.specs["k_pivot"])
    return out
