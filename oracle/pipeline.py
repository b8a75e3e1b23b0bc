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
    pf = np.zeros((model.n_fluid, ns, nk)) if getattr(model, "n_fluid", 0) else None
    c = abi.AbcmbCosmo.from_buffer_copy(plan.bg.cosmo)
    lna = np.ascontiguousarray(plan.lna_save)
    c.lna_save = lna.ctypes.data
    c.k = kk.ctypes.data
    c.pt = pt.ctypes.data
    c.pt_fluid = pf.ctypes.data if pf is not None else None
    c.n_save, c.nk = ns, nk
    cos = (abi.AbcmbCosmo * 1)(c)
    ks = np.arange(nk) if k_subset is None else np.asarray(k_subset)
    items = (abi.AbcmbPertItem * ks.size)()
    for m, ik in enumerate(ks):
        items[m] = abi.AbcmbPertItem(k=kk[ik], lna_start=plan.lna_start[ik], rtol=plan.rtol[ik],
                                     atol=plan.atol[ik], cosmo=0, ik=int(ik))
    if getattr(model, "fluids", None) is not None:
        with O.fluids(model, plan.params):
            status, steps, states = O.pert(cos, items, model.settings, want_states=want_states)
    else:
        status, steps, states = O.pert(cos, items, model.settings, want_states=want_states)
    out = dict(pt=pt, pt_fluid=pf, status=status, steps=steps, states=states, cosmo=c, keep=(lna, kk))
    if k_subset is not None:
        return out
    out.update(downstream(model, plan, pt, lensing=lensing))
    return out


def downstream(model, plan, pt, lensing=None, chi_branch=None, cl_full=None):
    """Sources, LOS C_l, l-spline, lensing and P(k) of one cosmology from its table ``pt``.

    chi_branch: LOS distances that pick the Bessel branches (``oracle.los_cl_branch``)
    instead of the cosmology's own; cl_full: unlensed C_l on every l to lens (with
    C_l^phiphi of ``plan`` / ``pt``) instead of the LOS result.
    """
    kk = np.ascontiguousarray(plan.k if plan.k is not None else model.k_pt)
    ns = model.n_save
    pt = np.ascontiguousarray(pt)
    p = plan.params
    c = abi.AbcmbCosmo.from_buffer_copy(plan.bg.cosmo)
    lna = np.ascontiguousarray(plan.lna_save)
    c.lna_save, c.k, c.pt = lna.ctypes.data, kk.ctypes.data, pt.ctypes.data
    c.n_save, c.nk = ns, kk.size
    if cl_full is None:
        src = O.sources(pt, kk, ns, plan.bgv, plan.lna_save, model.k_transfer, model.scales)
        los_args = (src, plan.bgv[0], plan.bg.tau0, model.k_transfer, model.bes_tab, model.bes_rng, model.bes_l,
                    model.ell_idx, c.A_s, c.n_s, model.specs["k_pivot"])
        cl_raw = O.los_cl(*los_args) if chi_branch is None else O.los_cl_branch(*los_args, chi_branch)
        cl_full = cl_raw @ model.ell_op.T
        out = dict(src=src, cl_raw=cl_raw, cl_full=cl_full)
    else:
        out = dict(cl_full=cl_full)
    do_lens = model.specs["lensing"] if lensing is None else lensing
    if do_lens:
        clpp = O.clpp(c, pt, plan.lens_lna, plan.lens_aH, plan.lens_tau, plan.chi_rec,
                      model.lensing_ells.astype(np.float64))
        lensed = O.lens(cl_full, clpp, model.mu, model.w)
        out.update(clpp=clpp, lensed=lensed)
    out["pk"] = O.pk(pt, kk, ns, plan.lna_save, model.k_pk, c.A_s, c.n_s, model.specs["k_pivot"])
    return out


def _pert_inputs(model, plan, ks, pt):
    kk = np.ascontiguousarray(plan.k if plan.k is not None else model.k_pt)
    c = abi.AbcmbCosmo.from_buffer_copy(plan.bg.cosmo)
    lna = np.ascontiguousarray(plan.lna_save)
    c.lna_save, c.k, c.pt = lna.ctypes.data, kk.ctypes.data, pt.ctypes.data
    c.n_save, c.nk = model.n_save, kk.size
    cos = (abi.AbcmbCosmo * 1)(c)
    items = (abi.AbcmbPertItem * ks.size)()
    for m, ik in enumerate(ks):
        items[m] = abi.AbcmbPertItem(k=kk[ik], lna_start=plan.lna_start[ik], rtol=plan.rtol[ik],
                                     atol=plan.atol[ik], cosmo=0, ik=int(ik))
    return cos, items, (lna, kk)


def frozen_step_tangents(model, base, plus, minus, steps, k_subset, mode="chord"):
    """Derivatives of the perturbation table at the base plan's step sequence.

    Records every k-mode's steps and chord-iteration counts at ``base``,
    replays them at ``plus[d]`` / ``minus[d]`` (plans at theta +- steps[d]
    with frozen grids) and takes central differences.
    mode "chord": the recorded iteration counts with the base's chord
    matrices (the derivative a tangent-linear integrator computes);
    "own_chord": the recorded counts with each replay's own chord matrices;
    "converged": every stage solved to rounding (the derivative of the
    implicit Runge-Kutta map at the frozen steps).
    Returns (pt, dpt): [21][n_save][nk] and [len(plus)][21][n_save][nk],
    filled on k_subset.
    """
    kk = base.k if base.k is not None else model.k_pt
    ks = np.asarray(k_subset)
    nk, ns = kk.size, model.n_save
    cap = int(model.settings.max_steps)
    scripts = np.zeros((ks.size, cap, O.SCRIPT_W))
    lens = np.zeros(ks.size, np.int32)
    pt = np.zeros((abi.N_PT_FIELDS, ns, nk))
    cos, items, keep = _pert_inputs(model, base, ks, pt)
    st, _ = O.pert_scripted(cos, items, model.settings, 1, scripts, lens)
    rmode = 3 if mode == "converged" else 2
    jac = cos if mode == "chord" else None
    if np.any(st != 0):
        raise RuntimeError(f"oracle record failed: status {np.unique(st)}")
    dpt = np.zeros((len(plus), abi.N_PT_FIELDS, ns, nk))
    for d, (pp, pm, h) in enumerate(zip(plus, minus, steps)):
        tabs = []
        for pl in (pp, pm):
            t = np.zeros((abi.N_PT_FIELDS, ns, nk))
            c2, it2, keep2 = _pert_inputs(model, pl, ks, t)
            st2, _ = O.pert_scripted(c2, it2, model.settings, rmode, scripts, lens, jac_cosmos=jac)
            if np.any(st2 != 0):
                raise RuntimeError(f"oracle replay failed: status {np.unique(st2)}")
            tabs.append(t)
        dpt[d] = (tabs[0] - tabs[1]) / (2.0 * h)
    return pt, dpt


def tangent_tables(model, base, plus, minus, steps, k_subset):
    """Tangent-linear perturbation tables of ``base`` on ``k_subset`` (oracle).

    Same step sequence as the plain integration; each stage tangent solved
    to rounding; the right-hand side's parameter derivative by central
    differences between ``plus[d]`` and ``minus[d]`` (theta +- steps[d],
    grids frozen at ``base``).  Returns (pt [21][n_save][nk],
    dpt [nd][21][n_save][nk], steps per k of the subset).
    """
    kk = base.k if base.k is not None else model.k_pt
    ks = np.asarray(k_subset)
    nk, ns, nd = kk.size, model.n_save, len(plus)
    pt = np.zeros((abi.N_PT_FIELDS, ns, nk))
    cos, items, keep = _pert_inputs(model, base, ks, pt)
    pcs, mcs, keeps = [], [], []
    for pl_list, out in ((plus, pcs), (minus, mcs)):
        for pl in pl_list:
            c2, _, k2 = _pert_inputs(model, pl, ks, pt)
            out.append(c2[0])
            keeps.append(k2)
    parr = (abi.AbcmbCosmo * nd)(*pcs)
    marr = (abi.AbcmbCosmo * nd)(*mcs)
    dpt = np.zeros((nd, abi.N_PT_FIELDS, ns, nk))
    if getattr(model, "fluids", None) is not None:
        with O.fluids(model, base.params):
            st, nst = O.pert_tangent(cos, parr, marr, np.asarray(steps, dtype=np.float64), items, model.settings,
                                     dpt)
    else:
        st, nst = O.pert_tangent(cos, parr, marr, np.asarray(steps, dtype=np.float64), items, model.settings, dpt)
    if np.any(st != 0):
        raise RuntimeError(f"oracle tangent integration failed: status {np.unique(st)}")
    return pt, dpt, nst
