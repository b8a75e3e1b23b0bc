"""Lensed-C_l Gaussian likelihood and its gradient with respect to seven
cosmological parameters.

Parameters theta = (omega_b, omega_cdm, h, ln(1e10 A_s), n_s, tau_reio,
sum m_nu); the neutrino mass goes to one massive species (the reference's
single-species parameterisation, abcmb/main.py:398-400), the remaining
N_eff in massless species (abcmb/main.py:412-414).

Loss: L = -1/2 sum_{l=2}^{l_max} sum_{X in TT,TE,EE} (C_l^X - Chat_l^X)^2 / sigma_lX^2
with the mock Chat = C(theta_mock) (1 + eps), eps ~ N(0, sqrt(2/(2l+1)))
from a seed, and sigma_lX = sqrt(2/(2l+1)) |Chat_l^X|.

Gradients:
  - ``fd_gradients``: central finite differences (relative step 1e-4, 15
    model evaluations per gradient) with every discrete grid choice frozen at
    the base point; ``batched=True`` puts all evaluations of all cosmologies
    into one ODE launch, ``batched=False`` is the plain path (one cosmology
    per launch).
  - ``tangent_gradients``: one tangent-linear ODE launch for all cosmologies
    gives the perturbation tables and their exact derivatives along the five
    parameters that reach the perturbation equations (omega_b, omega_cdm, h,
    tau_reio, m_nu); the host inputs' derivatives come from central
    differences of the (cheap, deterministic) host producer.  By default the
    LOS is then differentiated in one tangent-linear pass (the pointwise
    derivative, what automatic differentiation of the reference gives) and
    the lensing is differenced along each direction; ``downstream="fd"``
    differences the whole downstream (sources, LOS, lensing) with the tables
    moved along their tangents instead, which also picks up the jumps where
    the LOS switches from the Bessel tables to their asymptotic form.
"""
from __future__ import annotations

import copy
import dataclasses
import math

import numpy as np

from . import abi

NAMES = ("omega_b", "omega_cdm", "h", "ln10As", "n_s", "tau_reion", "m_nu")
RANGES = np.array([(0.020, 0.024), (0.10, 0.14), (0.62, 0.74), (2.9, 3.2), (0.92, 1.0), (0.03, 0.09), (0.06, 0.3)])
FIDUCIAL = np.array([0.02237, 0.1200, 0.6736, math.log(1e10 * 2.1e-9), 0.9649, 0.0544, 0.06])
FD_REL_STEP = 1e-4
ODE_DIRS = (0, 1, 2, 5, 6)   # parameters that reach the perturbation equations


def theta_to_params(theta) -> dict:
    t = [float(v) for v in theta]
    return {"omega_b": t[0], "omega_cdm": t[1], "h": t[2], "A_s": math.exp(t[3]) * 1e-10, "n_s": t[4],
            "tau_reion": t[5], "N_nu_massive": 1, "m_nu_massive": t[6]}


def latin_hypercube(n: int, seed: int) -> np.ndarray:
    from scipy.stats import qmc
    u = qmc.LatinHypercube(d=7, seed=seed).random(n)
    return RANGES[:, 0] + u * (RANGES[:, 1] - RANGES[:, 0])


def uneven_k_grids(n: int, seed: int, lo=7.05e-6, hi=0.62):
    rng = np.random.default_rng(seed)
    return [np.geomspace(lo, hi, int(rng.integers(300, 801))) for _ in range(n)]


class LensedClLikelihood:
    """Gaussian likelihood of lensed TT/TE/EE against seeded mocks (one per slot)."""

    def __init__(self, model, cl_mock_base, n_mocks: int = 1, seed: int = 0):
        import torch
        self.model = model
        l = model.ells.astype(np.float64)
        sd = np.sqrt(2.0 / (2 * l + 1))
        rng = np.random.default_rng(seed)
        base = np.stack([np.asarray(c, dtype=np.float64) for c in cl_mock_base])  # [3][n_l]
        mocks = base[None] * (1 + rng.normal(size=(n_mocks, 3, l.size)) * sd)
        self.mock = torch.as_tensor(mocks, device="cuda")
        self.inv_var = torch.as_tensor(1.0 / (sd[None, None] * np.abs(mocks)) ** 2, device="cuda")

    def loglike(self, outs, mock_index):
        import torch
        cl = torch.stack([torch.stack([o.ClTT, o.ClTE, o.ClEE]) for o in outs])
        return self.loglike_cl(cl, mock_index)

    def loglike_cl(self, cl, mock_index):
        """L for spectra cl[B][3][n_l] (TT, TE, EE on l = 2..l_max) against mocks ``mock_index``."""
        import torch
        idx = torch.as_tensor(np.asarray(mock_index, dtype=np.int64)).pin_memory().to(cl.device, non_blocking=True)
        return -0.5 * torch.sum((cl - self.mock[idx]) ** 2 * self.inv_var[idx], dim=(1, 2))


def fd_gradients(model, like, thetas, k_grids=None, batched=True, mock_index=None):
    """Central-difference gradients d L / d theta for each row of ``thetas``."""
    thetas = np.atleast_2d(thetas)
    nc = thetas.shape[0]
    mock_index = list(range(nc)) if mock_index is None else list(mock_index)
    k_grids = k_grids or [None] * nc
    base_params = [model.derived(theta_to_params(t)) for t in thetas]
    base_plans = model.prepare_many(base_params, k_list=k_grids)
    pert_params, freeze, ks, slot = [], [], [], []
    steps = np.zeros_like(thetas)
    for c in range(nc):
        for p in range(7):
            h = FD_REL_STEP * abs(thetas[c, p])
            steps[c, p] = h
            for sgn in (+1, -1):
                t = thetas[c].copy()
                t[p] += sgn * h
                pert_params.append(model.derived(theta_to_params(t)))
                freeze.append(base_plans[c])
                ks.append(k_grids[c])
                slot.append(mock_index[c])
    pert_plans = model.prepare_many(pert_params, freeze_list=freeze, k_list=ks)
    all_plans = base_plans + pert_plans
    all_slots = mock_index + slot
    if batched:
        outs = model.execute(all_plans)
        L = like.loglike(outs, all_slots).cpu().numpy()
    else:
        L = np.zeros(len(all_plans))
        for i, pl in enumerate(all_plans):
            o = model.execute([pl])
            L[i] = float(like.loglike(o, [all_slots[i]])[0])
    Lb = L[:nc]
    Lp = L[nc:].reshape(nc, 7, 2)
    grad = (Lp[:, :, 0] - Lp[:, :, 1]) / (2 * steps)
    return Lb, grad


def with_primordial(plan, A_s: float, n_s: float):
    """Copy of ``plan`` with another primordial spectrum (shares every table)."""
    from . import abi
    bg = copy.copy(plan.bg)
    bg.cosmo = abi.AbcmbCosmo.from_buffer_copy(plan.bg.cosmo)
    bg.cosmo.A_s, bg.cosmo.n_s = A_s, n_s
    return dataclasses.replace(plan, params=dict(plan.params, A_s=A_s, n_s=n_s), bg=bg)


def _pm_params(model, theta, step):
    """Derived parameters at theta +- step along each ODE direction, in the order
    (direction, +), (direction, -)."""
    out = []
    for p in ODE_DIRS:
        for sgn in (+1, -1):
            t = theta.copy()
            t[p] += sgn * step[p]
            out.append(model.derived(theta_to_params(t)))
    return out


def _seed_inputs(model, thetas, base_plans, k_grids, steps):
    """Plans at theta +- step along the ODE directions (grids frozen at the base
    plans) and the tangent seeds of the host inputs from them."""
    from .model import tangent_seeds
    nc = thetas.shape[0]
    pm_params, freeze, ks = [], [], []
    for c in range(nc):
        q = _pm_params(model, thetas[c], steps[c])
        pm_params += q
        freeze += [base_plans[c]] * len(q)
        ks += [k_grids[c]] * len(q)
    nd = len(ODE_DIRS)
    if _forward_host(model):
        from concurrent.futures import ThreadPoolExecutor
        import os
        with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
            pm = list(ex.map(lambda c: model.tangent_plans(base_plans[c], pm_params[c * 2 * nd:(c + 1) * 2 * nd],
                                                           [steps[c, p] for p in ODE_DIRS]), range(nc)))
    else:
        pm = model.prepare_many(pm_params, freeze_list=freeze, k_list=ks)
        pm = [pm[c * 2 * nd:(c + 1) * 2 * nd] for c in range(nc)]
    seeds = [tangent_seeds(pm[c][0::2], pm[c][1::2], [steps[c, p] for p in ODE_DIRS]) for c in range(nc)]
    return pm, seeds


def _forward_host(model) -> bool:
    """Tangent seeds from the producer's forward mode (else, with
    host_tangents="fd", from frozen-grid central differences)."""
    return getattr(model, "host_tangents", "fd") == "forward"


def _tangent_inputs(model, thetas, k_grids, steps, threads: int = None):
    """Host side of the tangent gradients: base plans, plans at theta +- step along
    the ODE directions (grids frozen at the base), and the tangent seeds.  One
    thread pool: each cosmology's frozen plans are queued as soon as its base
    plan is ready, so no thread waits for the slowest base."""
    import os
    from concurrent.futures import ThreadPoolExecutor, as_completed
    from .model import tangent_seeds
    nc = thetas.shape[0]
    k_grids = list(k_grids) if k_grids is not None else [None] * nc
    threads = threads or min(16, os.cpu_count() or 1)
    base_plans, pm = [None] * nc, [None] * nc
    with ThreadPoolExecutor(threads) as ex:
        fb = {ex.submit(model.prepare, model.derived(theta_to_params(thetas[c])), None, k_grids[c]): c
              for c in range(nc)}
        fz = {}
        forward = _forward_host(model)
        for f in as_completed(fb):
            c = fb[f]
            base_plans[c] = f.result()
            q = _pm_params(model, thetas[c], steps[c])
            if forward:
                fz[c] = ex.submit(model.tangent_plans, base_plans[c], q, [steps[c, p] for p in ODE_DIRS])
            else:
                fz[c] = [ex.submit(model.prepare, qq, base_plans[c], k_grids[c]) for qq in q]
        for c in range(nc):
            pm[c] = fz[c].result() if forward else [f.result() for f in fz[c]]
    seeds = [tangent_seeds(pm[c][0::2], pm[c][1::2], [steps[c, p] for p in ODE_DIRS]) for c in range(nc)]
    if getattr(model, "device_background", False):
        # recombination-only tables: the GPU adds the reionization, its
        # implicit-function derivatives replace the x_e and z_reion seeds
        dtarget = np.zeros((nc, abi.N_TANGENT_DIRS))
        dtarget[:, ODE_DIRS.index(NAMES.index("tau_reion"))] = 1.0
        model.complete_backgrounds(base_plans, seeds, dtarget)
        model.complete_backgrounds([q for row in pm for q in row])
    return base_plans, pm, seeds



def _directional_losses(model, like, thetas, steps, base_plans, pm, pts_b, kdev_b, dpts, mock_index):
    """L at theta +- step along all seven directions ([nc][7][2], on the GPU): the
    downstream on tables moved along their tangents (ODE directions) or with
    another primordial spectrum (ln A_s, n_s)."""
    import torch
    from .model import _moved_tables
    nc = thetas.shape[0]
    per = 2 * 7
    group = max(1, model.chunk // per)
    _, _, stage = model._staging()
    hs_d = stage.add(np.stack([steps, -steps], axis=-1)[:, list(ODE_DIRS)])   # [nc][nd][2]
    stage.flush()
    Lp = []
    for c0 in range(0, nc, group):
        cs = range(c0, min(nc, c0 + group))
        plans, pts, kdev, slots = [], [], [], []
        for c in cs:
            pt = pts_b[c]
            kd = kdev_b[c]
            moved = _moved_tables(torch, pt, dpts[c], hs_d[c])   # pt + h dpt, all directions and signs
            for p in range(7):
                for si, sgn in enumerate((+1, -1)):
                    h = sgn * steps[c, p]
                    if p in ODE_DIRS:
                        d = ODE_DIRS.index(p)
                        plans.append(pm[c][2 * d + si])
                        pts.append(moved[2 * d + si])
                    else:
                        t = thetas[c].copy()
                        t[p] += h
                        q = theta_to_params(t)
                        plans.append(with_primordial(base_plans[c], q["A_s"], q["n_s"]))
                        pts.append(pt)
                    kdev.append(kd)
                    slots.append(mock_index[c])
        ds = model._downstream(plans, pts, kdev)
        Lp.append(like.loglike_cl(model.observed_cl(ds), slots).reshape(len(cs), 7, 2))
        del pts, moved
    return torch.cat(Lp)


def _directional_losses_linear(model, like, thetas, steps, base_plans, pm, pts, kdev, dpts, mock_index,
                               cl_out=None, jac_out=None):
    """L at theta ([nc]) and at theta +- step along all seven directions ([nc][7][2]),
    on the GPU, with the unlensed C_l of theta and their derivatives from the
    tangent-linear LOS (``Model.downstream_tangent``): the base C_l are its
    value, the moved C_l are C_l +- h dC_l; only C_l^phiphi and the lensing are
    evaluated at each point (the base one with theta's tables and plan).
    ``cl_out`` (a list, optional) receives the observed C_l at theta
    ([n][3][n_l] per group of cosmologies), ``jac_out`` their derivatives along
    the seven parameters ([n][7][3][n_l]: central differences of the lensed
    spectra between C_l +- h dC_l, the unlensed spectra moved along their exact
    derivatives)."""
    import torch
    nc = thetas.shape[0]
    nd = len(ODE_DIRS)
    n_l = model.ells.size
    plus = [pm[c][0::2] for c in range(nc)]
    minus = [pm[c][1::2] for c in range(nc)]
    sd = [[steps[c, p] for p in ODE_DIRS] for c in range(nc)]
    cl_t = model.downstream_tangent(base_plans, pts, kdev, dpts, plus, minus, sd)   # [nc][1+nd+2][3][nl]
    lensing = bool(model.specs["lensing"])
    from .model import _moved_tables
    _, _, stage = model._staging()
    # the column of cl_t that holds d C / d theta_p
    col = [1 + ODE_DIRS.index(p) if p in ODE_DIRS else (1 + nd if NAMES[p] == "ln10As" else 2 + nd)
           for p in range(7)]
    H = np.stack([steps, -steps], axis=-1)                       # [nc][7][2]: +h, -h
    H_d = stage.add(H)
    hs_d = stage.add(H[:, list(ODE_DIRS)])                       # [nc][nd][2]
    col_d = stage.add(np.array(col), torch.int64)
    stage.flush()
    # C(theta +- h e_p) = C + (+-h) dC/dp for every cosmology, direction and sign: [nc][7][2][3][nl]
    cls_all = torch.addcmul(cl_t[:, 0][:, None, None], H_d[:, :, :, None, None],
                            cl_t.index_select(1, col_d)[:, :, None])
    per = 1 + 2 * 7      # theta, then (direction, sign)
    group = max(1, model.chunk // per)
    L = []
    for c0 in range(0, nc, group):
        cs = range(c0, min(nc, c0 + group))
        plans, vpts, vk, slots = [], [], [], []
        for c in cs:
            # the lensing reads only delta_m, the tables' first field: move that alone
            moved = _moved_tables(torch, pts[c][:1], dpts[c][:, :1], hs_d[c]) if lensing else None
            slots += [mock_index[c]] * per
            if not lensing:
                continue
            plans.append(base_plans[c])
            vpts.append(pts[c])
            vk.append(kdev[c])
            for p in range(7):
                for si, sgn in enumerate((+1, -1)):
                    if p in ODE_DIRS:
                        d = ODE_DIRS.index(p)
                        plans.append(pm[c][2 * d + si])
                        vpts.append(moved[2 * d + si])
                    else:
                        t = thetas[c].copy()
                        t[p] += sgn * steps[c, p]
                        q = theta_to_params(t)
                        plans.append(with_primordial(base_plans[c], q["A_s"], q["n_s"]))
                        vpts.append(pts[c])
                    vk.append(kdev[c])
        n = len(cs)
        cl = torch.cat([cl_t[c0:c0 + n, 0][:, None], cls_all[c0:c0 + n].reshape(n, per - 1, 3, -1)], dim=1)
        cl = cl.reshape(n * per, 3, -1)
        if lensing:
            obs = model.lensed_from_cl(plans, vpts, vk, cl)[:, :3, :n_l]
        else:
            obs = cl[:, :, :n_l]
        L.append(like.loglike_cl(obs, slots).reshape(n, per))
        if cl_out is not None:
            cl_out.append(obs.reshape(n, per, 3, -1)[:, 0])
        if jac_out is not None:
            ob = obs.reshape(n, per, 3, -1)
            jac_out.append(torch.sub(ob[:, 1::2], ob[:, 2::2]).div_(2.0 * H_d[c0:c0 + n, :, 0, None, None]))
        del vpts, moved
    L = torch.cat(L)
    return L[:, 0], L[:, 1:].reshape(nc, 7, 2)


def tangent_gradients(model, like, thetas, k_grids=None, mock_index=None, phases=None, chunk: int = 64,
                      downstream: str = "tangent", cl_out=None, jac_out=None):
    """d L / d theta for each row of ``thetas`` from tangent-linear ODE launches.

    The cosmologies go through in chunks of ``chunk`` (one ODE launch each,
    the first three an eighth, a quarter and a half of that); a host thread
    prepares the next chunks' background tables and tangent seeds while the
    main thread enqueues the current chunk's ODE launch and downstream on the
    current stream without waiting for the GPU, so the launches run back to
    back.  ``cl_out`` (a list, optional) receives the observed C_l at each
    theta (tensors of [n][3][n_l], in order), ``jac_out`` (with the default
    downstream) their derivatives along the seven parameters ([n][7][3][n_l]).
    ``downstream``: "tangent"
    (default) differentiates the LOS in one tangent-linear pass (the Bessel
    lookups as interpolated segments and asymptotic forms, as automatic
    differentiation of the reference would) and differences only the
    lensing; "fd" differences the whole downstream (14 evaluations per
    cosmology, the same relative step as the plain finite-difference path,
    table/asymptotic jumps of the Bessel lookups included).  ``phases`` (optional
    dict) receives the main thread's time waiting for host inputs, enqueuing
    GPU work, and waiting at the end for the GPU to drain.
    """
    import time
    from concurrent.futures import ThreadPoolExecutor
    if jac_out is not None and downstream != "tangent":
        raise ValueError("jac_out needs the tangent-linear downstream")
    t_start = time.perf_counter()
    thetas = np.atleast_2d(thetas)
    nc = thetas.shape[0]
    mock_index = list(range(nc)) if mock_index is None else list(mock_index)
    k_grids = list(k_grids) if k_grids is not None else [None] * nc
    steps = FD_REL_STEP * np.abs(thetas)
    # doubling first chunks (an eighth, a quarter, a half of ``chunk``), so
    # the GPU starts early and each chunk's host preparation (about half the
    # GPU time of a chunk of the same size) ends about when the GPU has
    # finished the chunk before it
    chunk = max(1, chunk)
    bounds, c0 = [], 0
    for size in [max(1, chunk // 8), max(1, chunk // 4), max(1, chunk // 2)] + [chunk] * nc:
        if c0 >= nc:
            break
        bounds.append((c0, min(nc, c0 + size)))
        c0 += size
    import torch
    # Everything goes on the current stream without a host wait: chunk i's ODE
    # launch, its downstream and its losses are enqueued while the GPU still
    # works on earlier chunks (the staging copies are asynchronous), so the
    # GPU runs the launches back to back; the losses come back and the
    # launches' status is checked once at the end.
    Lb_d, Lp_d, handles = [], [], []
    t_host = t_enq = 0.0
    with ThreadPoolExecutor(1) as host:
        futs = [host.submit(_tangent_inputs, model, thetas[a:b], k_grids[a:b], steps[a:b]) for a, b in bounds]

        def launch(i):
            nonlocal t_host
            t0 = time.perf_counter()
            base_plans, pm, seeds = futs[i].result()
            t_host += time.perf_counter() - t0
            return base_plans, pm, model.launch_tangent(base_plans, seeds)

        cur = launch(0)
        for i, (a, b) in enumerate(bounds):
            # the next chunk's ODE launch goes into the stream ahead of this
            # chunk's downstream, whose many small launches the host enqueues
            # while the GPU runs it
            nxt = launch(i + 1) if i + 1 < len(bounds) else None
            t1 = time.perf_counter()
            base_plans, pm, h = cur
            run = model._ode_result(h)
            args = (model, like, thetas[a:b], steps[a:b], base_plans, pm, run["pts"], run["kdev"], run["dpts"],
                    mock_index[a:b])
            if downstream == "tangent":
                # the tangent-linear LOS also gives theta's own C_l: L(theta) is
                # one more lensing variant, no separate base downstream
                lb, lp = _directional_losses_linear(*args, cl_out=cl_out, jac_out=jac_out)
            else:
                ds = model._downstream(base_plans, run["pts"], run["kdev"])
                obs = model.observed_cl(ds)
                if cl_out is not None:
                    cl_out.append(obs)
                lb = like.loglike_cl(obs, mock_index[a:b])
                lp = _directional_losses(*args)
                del ds
            Lb_d.append(lb)
            Lp_d.append(lp)
            # keep only what the status check reads: the tables and staging
            # buffers go back to the allocator (their reuse is ordered after
            # this chunk's work on the same stream)
            handles.append({key: h[key] for key in ("done", "flat", "allk", "status", "steps", "iters", "ev",
                                                     "tangent")})
            del run, args, h, cur
            cur = nxt
            t_enq += time.perf_counter() - t1
    t0 = time.perf_counter()
    Lb = torch.cat(Lb_d).cpu().numpy()
    Lp = torch.cat(Lp_d).cpu().numpy()
    for h in handles:
        model._check_ode(h)
    t_drain = time.perf_counter() - t0
    grad = (Lp[:, :, 0] - Lp[:, :, 1]) / (2 * steps)
    if phases is not None:
        phases.update(wait_inputs_s=t_host, enqueue_s=t_enq, gpu_drain_s=t_drain,
                      total_s=time.perf_counter() - t_start, chunk=chunk)
    return Lb, grad


def _clpp_directional(model, thetas, steps, base_plans, pm, pts, kdev):
    """d C_l^phiphi / d theta_p with the perturbation tables held at theta's
    ([nc][7][n_l]): the Limber integral's host inputs (distances, aH, the
    cosmology fields, the primordial spectrum) differenced between the plans
    at theta +- step, on theta's own delta_m table."""
    import torch
    nc = thetas.shape[0]
    plans, tabs, ks = [], [], []
    for c in range(nc):
        for p in range(7):
            for si, sgn in enumerate((+1, -1)):
                if p in ODE_DIRS:
                    plans.append(pm[c][2 * ODE_DIRS.index(p) + si])
                else:
                    t = thetas[c].copy()
                    t[p] += sgn * steps[c, p]
                    q = theta_to_params(t)
                    plans.append(with_primordial(base_plans[c], q["A_s"], q["n_s"]))
                tabs.append(pts[c])
                ks.append(kdev[c])
    clpp = model.clpp_of(plans, tabs, ks).view(nc, 7, 2, -1)
    h2 = torch.as_tensor(2.0 * steps, dtype=torch.float64, device=clpp.device)
    return (clpp[:, :, 0] - clpp[:, :, 1]) / h2[:, :, None]


def adjoint_forward(model, thetas, k_grids=None):
    """The forward half of the reverse-mode gradient: plans, one primal ODE launch
    and the downstream of every row of ``thetas``.  Returns a state for
    ``adjoint_backward`` and the observed spectra [n][3][n_l]."""
    if getattr(model, "device_background", False):
        raise NotImplementedError("the reverse-mode gradient takes the host background (device_background=False)")
    if getattr(model, "fluids", None) is not None:
        raise NotImplementedError("the reverse-mode gradient covers the standard species; a model with user "
                                  "fluids takes forward-mode gradients (tangent_gradients, lensed_cl(mode='jacobian'))")
    thetas = np.atleast_2d(np.asarray(thetas, dtype=np.float64))
    k_grids = list(k_grids) if k_grids is not None else [None] * thetas.shape[0]
    plans = model.prepare_many([model.derived(theta_to_params(t)) for t in thetas], k_list=k_grids)
    run = model._run_ode(plans)
    ds = model._downstream(plans, run["pts"], run["kdev"])
    state = dict(thetas=thetas, k_grids=k_grids, plans=plans, run=run, ds=ds)
    return state, model.observed_cl(ds)


def adjoint_backward(model, state, obs_bar, cotangents_out=None):
    """d L / d theta ([n][7]) for the cotangent ``obs_bar`` ([n][3][n_l]) of the
    observed spectra of ``adjoint_forward``: the downstream backwards
    (``Model.downstream_adjoint``) to the cotangents of the perturbation
    tables, the discrete-adjoint ODE launch (``Model.ode_adjoint``) to those of
    the host inputs (cosmology fields, tau / x_e / T_m tables), then their
    contraction with the host inputs' derivatives along the parameters: the
    tangent seeds of the producer, the LOS background vectors and the Limber
    inputs of the plans at theta +- step (central differences of the host
    producer, as the tangent-linear path seeds them).  ``cotangents_out``
    (optional list) receives per cosmology the host-input cotangents before
    that contraction (dict: par, tau, xe, tm, bgv, clpp)."""
    thetas, plans, run, ds = state["thetas"], state["plans"], state["run"], state["ds"]
    n = thetas.shape[0]
    steps = FD_REL_STEP * np.abs(thetas)
    pm, seeds = _seed_inputs(model, thetas, plans, state["k_grids"], steps)
    adj = model.downstream_adjoint(plans, run["pts"], run["kdev"], ds, obs_bar)
    cot = model.ode_adjoint(plans, adj["pt_bar"])
    from . import abi
    names = [f[0] for f in abi.AbcmbCosmo._fields_]
    f_tau0, f_as, f_ns = names.index("tau0"), names.index("A_s"), names.index("n_s")
    lp = np.zeros((n, 7))
    if model.specs["lensing"]:
        import torch
        clpp_d = _clpp_directional(model, thetas, steps, plans, pm, run["pts"], run["kdev"])
        lp = torch.einsum("cl,cpl->cp", adj["clpp_bar"], clpp_d).cpu().numpy()
    bgv_bar = adj["bgv_bar"].cpu().numpy()
    scal = adj["scal_bar"].cpu().numpy()
    g = np.zeros((n, 7))
    for i in range(n):
        par = cot[i]["par"].cpu().numpy().copy()
        par[f_tau0] += scal[i, 2]
        par[f_as] += scal[i, 0]
        par[f_ns] += scal[i, 1]
        tau, xe, tm = (cot[i][key].cpu().numpy() for key in ("tau", "xe", "tm"))
        if cotangents_out is not None:
            cotangents_out.append(dict(par=par, tau=tau, xe=xe, tm=tm, bgv=bgv_bar[i],
                                       clpp=adj["clpp_bar"][i].cpu().numpy() if adj["clpp_bar"] is not None
                                       else None))
        sd = seeds[i]
        for p in range(7):
            if p in ODE_DIRS:
                d = ODE_DIRS.index(p)
                dbgv = (pm[i][2 * d].bgv - pm[i][2 * d + 1].bgv) / (2.0 * steps[i, p])
                g[i, p] = (par @ sd.par[d] + tau @ sd.tau[d] + xe @ sd.xe[d] + tm @ sd.tm[d] +
                           np.sum(bgv_bar[i] * dbgv))
            elif NAMES[p] == "ln10As":
                g[i, p] = par[f_as] * plans[i].bg.cosmo.A_s
            else:
                g[i, p] = par[f_ns]
        g[i] += lp[i]
    return g


def adjoint_gradients(model, like, thetas, k_grids=None, mock_index=None, chunk: int = 64, phases=None,
                      cl_out=None):
    """d L / d theta for each row of ``thetas`` by reverse mode on the GPU.

    Per chunk of ``chunk`` cosmologies: ``adjoint_forward`` (one primal ODE
    launch and the downstream) gives L; the cotangent of the observed spectra
    (torch autograd through the likelihood) goes through ``adjoint_backward``:
    the lensing, l-spline, sources/LOS/k-quadrature and Limber adjoints, one
    discrete-adjoint ODE launch, and the contraction of the host inputs'
    cotangents with their derivatives along the parameters -- the reference's
    jax.vjp through the background and recombination, here by the host's
    directional derivatives.  ``cl_out`` (optional list) receives the
    observed spectra of each chunk; ``phases`` (optional dict) the time spent
    in the forward and backward halves.
    """
    import time
    import torch
    t_start = time.perf_counter()
    thetas = np.atleast_2d(thetas)
    nc = thetas.shape[0]
    mock_index = list(range(nc)) if mock_index is None else list(mock_index)
    k_grids = list(k_grids) if k_grids is not None else [None] * nc
    Lb, grads = [], []
    t_fwd = t_bwd = 0.0
    for a in range(0, nc, max(1, chunk)):
        b = min(nc, a + max(1, chunk))
        t0 = time.perf_counter()
        state, obs = adjoint_forward(model, thetas[a:b], k_grids[a:b])
        obs = obs.detach().clone().requires_grad_(True)
        L = like.loglike_cl(obs, mock_index[a:b])
        (obs_bar,) = torch.autograd.grad(L.sum(), obs)
        if cl_out is not None:
            cl_out.append(obs.detach())
        t1 = time.perf_counter()
        grads.append(adjoint_backward(model, state, obs_bar))
        Lb.append(L.detach().cpu().numpy())
        t_fwd += t1 - t0
        t_bwd += time.perf_counter() - t1
        del state
    if phases is not None:
        phases.update(forward_s=t_fwd, backward_s=t_bwd, total_s=time.perf_counter() - t_start, chunk=chunk)
    return np.concatenate(Lb), np.concatenate(grads)
