"""Benchmark: gradient evaluations per second of the 7-parameter lensed-C_l
likelihood (l <= 2500) on one MI355X.

Workload (batched gradients): ``--batch`` (64) Latin-hypercube cosmologies
(omega_b, omega_cdm, h, ln 1e10 A_s, n_s, tau_reio, sum m_nu over the stated
ranges, seed 0), each with its own uneven log-k grid of 300-800 modes in
[7.05e-6, 0.62] Mpc^-1 (seed 0), one massive neutrino species (3 momentum
bins x l <= 17, state size 100), lensed TT/TE/EE to l = 2500 in FP64 (ODE
and LOS), Gaussian log-likelihood against a per-cosmology seeded mock.  The
ODE runs at the reference's tolerances with the large-k ones tightened
(``--tolerances accurate``, the default) so that the gradients are within
1e-3 of their converged values; the same batch at the reference's own
tolerances and a ``--batch-large`` (256) batch are timed once beside it.
One step = the gradients of all ``--batch`` cosmologies: tangent-linear ODE
launches of ``--chunk`` cosmologies each (perturbation tables and their
derivatives along omega_b, omega_cdm, h, tau_reio, m_nu), each followed by
a tangent-linear LOS pass and the lensing differenced along all seven
directions (``--fd-downstream``: the whole downstream differenced), with the
next chunk's host preparation overlapping the GPU work (``--gradient fd`` instead:
central finite differences, 15 full model evaluations per gradient, all in
one ODE launch).

Also reported: the launch shapes the autotuner picked (untimed, at setup), the
reverse-mode gradient of the same batch (``adjoint_gradients``: downstream and
ODE adjoints, one step; ``--no-adjoint`` skips it), the
plain path (one cosmology per launch, FP64, central finite differences, the
reference's tolerances) on a few of the same cosmologies, the gradient errors
of the timed path, of the reference-tolerance path and of the plain path
against converged gradients, the roofline of the dominant kernel (the perturbation
ODE, FP64 FLOP/s against the FP64 peak, timed with HIP events on its stream),
and the CPU oracle on a bounded sample of the same work.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grad-evals/s, 7-param lensed Cl ℓ≤2500, vs plain GPU path at ≤1e-4 Cl err"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0
# algorithmic FP64 flops of the tangent-linear LOS per (cosmology, l, k, tau) node:
# value 4 sources x 1 multiply-add = 8; per direction the source derivatives and the
# Bessel slopes times the distance derivative, 4 x 2 multiply-adds + 1 = 17
LOS_T_FLOPS_VALUE, LOS_T_FLOPS_PER_DIR = 8, 17
# algorithmic FP64 flops of the lensing per (cosmology, mu, l): Wigner recursions
# (4 flops a step; 2 in the Cgl pass, 12 in the correlation pass, 4 in the
# back-transform) 72, Cgl sums 4, X functions 14, correlation sums 40,
# back-transform sums 8
LENS_FLOPS_PER_MU_L = 138


def ode_flops(model, steps: int, iters: int) -> float:
    """Algorithmic FP64 flops of the ODE kernel for the given step / chord-iteration counts.

    Per step: block Gauss-Jordan 2 b^3 per hierarchy block, Woodbury setup
    4 b per row, stage bases and error estimate (6*7 + 16) per row.
    Per chord iteration: right-hand side (5 terms, two metric sums) 14, residual
    3, block solve 2 b + Woodbury 4, norm 4 flops per row.
    """
    lay = model.layout
    blocks = [1, 1, lay.i_ur - 2, lay.l_ur + 1] + [lay.l_ncdm + 1] * lay.n_q
    rows_b = sum(b * b for b in blocks)          # sum over rows of their block size
    f_step = sum(2 * b ** 3 for b in blocks) + 4 * rows_b + 58 * lay.n
    f_iter = 25 * lay.n + 2 * rows_b
    return steps * f_step + iters * f_iter


def ode_flops_tangent(model, steps: int, iters: int, n_dir: int = 5) -> float:
    """Algorithmic FP64 flops of the tangent-linear ODE kernel.

    The value part costs what the primal costs; each tangent direction adds
    per chord iteration: right-hand side with the scalar derivatives (5 flops
    per term, 8 for the metric sums, 4 for the couplings) 37, residual 3,
    block solve 2 b + Woodbury 10, update 1, convergence sums 6 per row; per
    step the stage bases and predictors, 50 per row.
    """
    lay = model.layout
    blocks = [1, 1, lay.i_ur - 2, lay.l_ur + 1] + [lay.l_ncdm + 1] * lay.n_q
    rows_b = sum(b * b for b in blocks)
    per_dir = iters * (57 * lay.n + 2 * rows_b) + steps * 50 * lay.n
    return ode_flops(model, steps, iters) + n_dir * per_dir


def kernel_instance(model, n_dir: int) -> str:
    """The ODE kernel template instance the launch selects for this layout (threads per
    workgroup by state size, block bound by the largest hierarchy block)."""
    lay = model.layout
    maxb = max(lay.i_ur - 2, lay.l_ur + 1, lay.l_ncdm + 1 if lay.n_q else 1)
    bmax = 26 if maxb <= 26 else 32 if maxb <= 32 else 64
    nt = 64 if lay.n <= 64 else 128 if lay.n <= 128 else 192
    return f"pert_fwd_kernel<{nt}, {bmax}, {n_dir}>"


# ODE tolerances (the reference's are rtol 1e-5 / 1e-4 and atol 1e-10 / 1e-6 below /
# above k = 0.01 Mpc^-1, abcmb/model_specs.py:63-66).  "accurate": the large-k
# rtol / 100 and atol / 5, the cheapest setting measured to bring the
# likelihood gradient within 1e-3 of its converged value (the large-k atol is
# the expensive part); "converged": large-k tolerances / 300 and small-k / 30,
# the reference the gradient errors are measured against.
TOL_ACCURATE = {"rtol_large_k_PE": 1.0e-6, "atol_large_k_PE": 2.0e-7}
TOL_CONVERGED = {"rtol_large_k_PE": 1e-4 / 300, "atol_large_k_PE": 1e-6 / 300, "rtol_small_k_PE": 1e-5 / 30,
                 "atol_small_k_PE": 1e-10 / 30}


def build_model(tol: dict = None, los_precision: str = "fp64"):
    """The benchmark's model, at the reference's ODE tolerances unless ``tol``
    overrides some of them."""
    from abcmb_amd.model import Model, MassiveNeutrino
    # max_steps_PE raised from the reference default 2048: with massive neutrinos the
    # highest modes (k ~ 0.6 Mpc^-1) need up to ~2000 steps in this integrator
    return Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, los_precision=los_precision,
                 max_steps_PE=4096 if not tol else 32768, **(tol or {}))


def cl_errors(gpu, ref, n_l):
    """Max error of lensed TT/TE/EE/BB ([4][>= n_l] arrays, l = 2..) against ``ref``:
    relative per l for TT, EE, BB, relative to max |TE| for TE."""
    out = {}
    for x, name in enumerate(("TT", "TE", "EE", "BB")[:min(len(gpu), len(ref))]):
        g, r = np.asarray(gpu[x][:n_l]), np.asarray(ref[x][:n_l])
        den = np.max(np.abs(r)) if name == "TE" else np.abs(r)
        out[name] = float(np.max(np.abs(g - r) / den))
    return out


def cpu_baseline(model, plan, gpu_lensed):
    """The oracle (CPU FP64 restatement of the whole path) on one cosmology of the
    workload: one complete forward evaluation (every k-mode, sources, LOS,
    lensing) on all host cores, and on one core a bounded sample (16 of the
    k-modes, scaled by the k count, plus the whole downstream).  A central
    difference gradient costs 15 forward evaluations.  Also returns the GPU's
    lensed C_l error against the complete oracle evaluation."""
    from oracle import oracle as O
    from oracle import pipeline
    threads = O.max_threads()
    nk = plan.k.size
    t0 = time.perf_counter()
    o = pipeline.run(model, plan)
    t_all = time.perf_counter() - t0
    if np.any(o["status"] != 0):
        raise RuntimeError("oracle integration failed on the baseline cosmology")
    sub = np.linspace(0, nk - 1, 16).astype(int)
    O.set_threads(1)
    try:
        t0 = time.perf_counter()
        pipeline.run(model, plan, k_subset=sub)
        t_ode1 = (time.perf_counter() - t0) * nk / sub.size
        t0 = time.perf_counter()
        pipeline.downstream(model, plan, o["pt"])
        t_down1 = time.perf_counter() - t0
    finally:
        O.set_threads(threads)
    err = cl_errors(gpu_lensed, o["lensed"], model.ells.size)
    base = {"value": 1.0 / (15 * t_all), "unit": "grad-evals/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(),
            "sample": (f"one complete oracle forward evaluation of cosmology 0 ({nk} k-modes, sources, LOS, "
                       f"lensing to l=2500) on {threads} threads: {t_all:.1f} s; x15 evaluations per "
                       f"central-difference gradient"),
            "one_core": {"value": 1.0 / (15 * (t_ode1 + t_down1)), "cores": 1,
                         "sample": (f"oracle ODE on {sub.size} of {nk} k-modes on 1 thread, scaled by the k count "
                                    f"({t_ode1:.1f} s), + the whole downstream on 1 thread ({t_down1:.1f} s); x15")}}
    return base, err


def copy_calibration(nbytes: int):
    """Five copies of ``nbytes`` between two device buffers (FETCH_SIZE / WRITE_SIZE reference)."""
    import torch
    n = nbytes // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": n * 8, "launches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--batch", type=int, default=64,
                    help="cosmologies per step (the gradient target's 64-cosmology batch)")
    ap.add_argument("--batch-large", type=int, default=256,
                    help="cosmologies of the larger batched-gradient workload timed once beside the headline "
                         "(0: skip)")
    ap.add_argument("--tolerances", choices=("accurate", "reference"), default="accurate",
                    help="ODE tolerances of the timed path: the reference's, or the large-k ones tightened "
                         "until the gradient is within 1e-3 of its converged value")
    ap.add_argument("--plain", type=int, default=2, help="cosmologies timed on the plain path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--chunk", type=int, default=64, help="cosmologies per tangent-linear ODE launch")
    ap.add_argument("--gradient", choices=("tangent", "fd"), default="tangent",
                    help="batched gradient method (the plain path is always finite differences)")
    ap.add_argument("--fd-downstream", action="store_true",
                    help="difference the whole downstream along each direction instead of differentiating the "
                         "LOS in one tangent-linear pass")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the one-step timing of the reverse-mode (adjoint) gradient beside the headline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (5 times): the reference for "
                         "scaling HBM byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; sharding the cosmologies across GPUs "
                                   "is designed but not launched"}))
        sys.exit(2)
    import torch
    from abcmb_amd import autotune
    from abcmb_amd import likelihood as LK

    t_setup = time.perf_counter()
    tol = TOL_ACCURATE if args.tolerances == "accurate" else None
    model = build_model(tol)
    mref = build_model() if tol else model          # the reference's tolerances
    n_all = max(args.batch, args.batch_large)
    thetas_all = LK.latin_hypercube(n_all, seed=0)
    k_all = LK.uneven_k_grids(n_all, seed=0)
    thetas, k_grids = thetas_all[:args.batch], k_all[:args.batch]
    # mock: C_l at the fiducial point, noise per cosmology slot
    fid = model.execute([model.prepare(model.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(model, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=n_all, seed=1)
    tuned = None
    if not args.no_autotune and args.gradient == "tangent":
        # launch shapes picked on a 16-cosmology sample of the workload (untimed)
        ns = min(16, args.batch)
        bp, _, seeds = LK._tangent_inputs(model, thetas[:ns], k_grids[:ns], LK.FD_REL_STEP * np.abs(thetas[:ns]))
        tuned = autotune.tune(model, bp, seeds)
        if mref is not model:
            autotune.apply(mref, tuned)
        del bp, seeds
        print(f"# autotune {tuned}", file=sys.stderr, flush=True)
    print(f"# setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    phases = {}

    def grads(m, th, kg, ph=None, chunk=None):
        if args.gradient == "tangent":
            return LK.tangent_gradients(m, like, th, k_grids=kg, phases=ph, chunk=chunk or args.chunk,
                                        downstream="fd" if args.fd_downstream else "tangent")
        return LK.fd_gradients(m, like, th, k_grids=kg, batched=True)

    for w in range(args.warmup):
        grads(model, thetas, k_grids)
        print(f"# warmup {w} done", file=sys.stderr, flush=True)

    model.timing = []
    model.call_timing = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        L, grad = grads(model, thetas, k_grids, phases)
        print(f"# step {s} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timing = model.timing
    call_timing = model.call_timing
    model.timing = model.call_timing = None
    value = args.batch * args.steps / dt

    def timed_once(m, n):
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, g = grads(m, thetas_all[:n], k_all[:n])
        torch.cuda.synchronize()
        return n / (time.perf_counter() - t), g

    # the same batch at the reference's tolerances, and the larger batch (one step each)
    ref_rate, g_ref = timed_once(mref, args.batch) if mref is not model else (value, grad)
    print(f"# reference tolerances {ref_rate:.2f} grad-evals/s", file=sys.stderr, flush=True)
    large = None
    if args.batch_large:
        large = {"cosmologies": args.batch_large, "steps": 1}
        large["grad_evals_per_s"], _ = timed_once(model, args.batch_large)
        if mref is not model:
            large["reference_tolerances_grad_evals_per_s"], _ = timed_once(mref, args.batch_large)
        print(f"# {args.batch_large} cosmologies {large}", file=sys.stderr, flush=True)

    # plain path: one cosmology per launch, central finite differences, the reference's tolerances
    np_ = args.plain
    t0 = time.perf_counter()
    Lp, gp = LK.fd_gradients(mref, like, thetas[:np_], k_grids=k_grids[:np_], batched=False)
    torch.cuda.synchronize()
    plain = np_ / (time.perf_counter() - t0)
    gb = grad[:np_]

    def err_to_norm(g, ref):
        """max over parameters of |g - ref| relative to ref's norm, each component in
        units of its parameter's magnitude (components near zero stay meaningful)"""
        scale = np.linalg.norm(ref * thetas[:np_], axis=1, keepdims=True) / np.abs(thetas[:np_])
        return float(np.max(np.abs(g - ref) / scale))

    print(f"# plain path {plain:.4f} grad-evals/s", file=sys.stderr, flush=True)

    # accuracy on the plain cosmologies: C_l and L of the timed path against one-cosmology
    # launches of the same model, gradients against a converged FP64 reference (the
    # tangent-linear path at far tighter ODE tolerances)
    cl_b = []
    LK.tangent_gradients(model, like, thetas[:np_], k_grids=k_grids[:np_], chunk=np_, cl_out=cl_b)
    cl_b = torch.cat(cl_b).cpu().numpy()
    plans_p = model.prepare_many([model.derived(LK.theta_to_params(t)) for t in thetas[:np_]], k_list=k_grids[:np_])
    outs_p = [model.execute([pl])[0] for pl in plans_p]
    cl_p = np.stack([np.stack([o.ClTT.cpu().numpy(), o.ClTE.cpu().numpy(), o.ClEE.cpu().numpy()]) for o in outs_p])
    cl_diff_plain = max(max(cl_errors(cl_b[i], cl_p[i], model.ells.size)[x] for x in ("TT", "TE", "EE"))
                        for i in range(np_))
    mconv = build_model(TOL_CONVERGED)
    if tuned:
        autotune.apply(mconv, tuned)
    _, g_conv = LK.tangent_gradients(mconv, like, thetas[:np_], k_grids=k_grids[:np_], chunk=np_)
    del mconv
    grad_err_conv = err_to_norm(gb, g_conv)
    ref_err_conv = err_to_norm(g_ref[:np_], g_conv)
    plain_err_conv = err_to_norm(gp, g_conv)
    agree = err_to_norm(gb, gp)
    agree_rel = float(np.max(np.abs(gp - gb) / (np.abs(gp) + 1e-300)))
    print(f"# gradient error vs converged: timed {grad_err_conv:.2e}, reference tolerances {ref_err_conv:.2e}, "
          f"plain {plain_err_conv:.2e}", file=sys.stderr, flush=True)

    # the reverse-mode gradient (downstream and ODE adjoints) on the same batch, one step:
    # rate, the adjoint launch's time, agreement with the timed path and error vs converged
    adjoint = None
    if args.gradient == "tangent" and args.adjoint:
        model.timing = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ph_a = {}
        _, g_adj = LK.adjoint_gradients(model, like, thetas, k_grids=k_grids, chunk=args.chunk, phases=ph_a)
        torch.cuda.synchronize()
        rate_a = args.batch / (time.perf_counter() - t0)
        t_adj = [t for t in model.timing if t["kernel"] == "adjoint"]
        t_pri = [t for t in model.timing if t["kernel"] == "primal"]
        model.timing = None
        scale_b = np.linalg.norm(grad * thetas, axis=1, keepdims=True) / np.abs(thetas)
        adjoint = {"grad_evals_per_s": rate_a, "speedup_vs_plain": rate_a / plain,
                   "ode_adjoint_ms_per_launch": sum(t["ms"] for t in t_adj) / max(1, len(t_adj)),
                   "ode_primal_ms_per_launch": sum(t["ms"] for t in t_pri) / max(1, len(t_pri)),
                   "launches": len(t_adj),
                   "grad_diff_vs_timed_rel_to_norm": float(np.max(np.abs(g_adj - grad) / scale_b)),
                   "grad_diff_vs_timed_median": float(np.median(np.max(np.abs(g_adj - grad) / scale_b, axis=1))),
                   "grad_err_vs_converged": err_to_norm(g_adj[:np_], g_conv),
                   "phases_s": ph_a,
                   "method": "one primal ODE launch + downstream, then the lensing / l-spline / LOS / Limber "
                             "adjoints and one discrete-adjoint ODE launch (checkpoints every 64 accepted "
                             "steps, segments recomputed, transposed stage solves), contracted with the host "
                             "seeds"}
        print(f"# adjoint {adjoint}", file=sys.stderr, flush=True)

    # the reduced-precision LOS contractions in the same gradient path (one step of the
    # same batch each): rate, and spectra and gradients against the FP64 LOS
    los_variants = {}
    if args.gradient == "tangent" and not args.fd_downstream:
        scale_all = np.linalg.norm(grad * thetas, axis=1, keepdims=True) / np.abs(thetas)
        for prec in ("fp32", "bf16"):
            mv = build_model(tol, los_precision=prec)
            if tuned:
                autotune.apply(mv, tuned)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            cl_v = []
            _, gv = LK.tangent_gradients(mv, like, thetas, k_grids=k_grids, chunk=args.chunk, cl_out=cl_v)
            torch.cuda.synchronize()
            rate_v = args.batch / (time.perf_counter() - t0)
            cl_v = torch.cat(cl_v).cpu().numpy()[:np_]
            los_variants[prec] = {
                "grad_evals_per_s": rate_v,
                "grad_diff_vs_fp64_los": float(np.max(np.abs(gv - grad) / scale_all)),
                "grad_err_vs_converged": err_to_norm(gv[:np_], g_conv),
                "cl_diff_vs_fp64_los": {x: max(cl_errors(cl_v[i], cl_b[i], model.ells.size)[x] for i in range(np_))
                                        for x in ("TT", "TE", "EE")}}
            del mv
        print(f"# LOS variants {los_variants}", file=sys.stderr, flush=True)

    ms = sum(t["ms"] for t in timing) / len(timing)
    if args.gradient == "tangent":
        flops = sum(ode_flops_tangent(model, t["steps"], t["iters"]) for t in timing) / len(timing)
        kname = kernel_instance(model, 5) + " (tangent-linear; FP64 VALU; peak = MI355X FP64 vector/matrix)"
    else:
        flops = sum(ode_flops(model, t["steps"], t["iters"]) for t in timing) / len(timing)
        kname = kernel_instance(model, 0) + " (FP64 VALU; peak = MI355X FP64 vector/matrix)"
    achieved = flops / (ms * 1e-3) / 1e12
    # HBM bytes per launch of the tangent kernel, from the two PMC runs summarised in profiles/
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pert_tangent_traffic.json")
    if args.gradient == "tangent" and os.path.exists(tfile):
        t = json.load(open(tfile))
        if (t.get("tolerances", "reference"), t.get("batch"), t.get("chunk")) == (args.tolerances, args.batch,
                                                                                   args.chunk):
            traffic = t["fetch_bytes_per_launch"] + t["write_bytes_per_launch"]
    # algorithmic HBM bytes of an ODE launch: the perturbation table it produces
    # (21 fields x saved points x k-modes, FP64, value and 5 tangents)
    ncomp = 1 + (5 if args.gradient == "tangent" else 0)
    table_bytes = sum(t["items"] for t in timing) / len(timing) * model.n_save * 21 * 8 * ncomp
    roof = {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic, "kernel": kname,
            "launch_ms": ms, "flops_per_launch": flops, "launches_per_step": len(timing) / args.steps,
            "chord_iters_per_step": sum(t["iters"] for t in timing) / max(1, sum(t["steps"] for t in timing)),
            "traffic_source": "profiles/pert_tangent_traffic.json (FETCH_SIZE + WRITE_SIZE per launch of the same "
                              "workload, scaled by a plain copy)" if traffic else None,
            "hbm": {"algorithmic_bytes_per_launch": table_bytes,
                    "algorithmic_GBps": table_bytes / (ms * 1e-3) / 1e9,
                    "counter_GBps": traffic / (ms * 1e-3) / 1e9 if traffic else None,
                    "peak_GBps": HBM_PEAK_GBPS,
                    "frac": table_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS}}
    others = []
    for name in ("abcmb_los_tangent", "lensing"):
        calls = call_timing.get(name, [])
        if not calls:
            continue
        cms = [e0.elapsed_time(e1) for e0, e1, _ in calls]
        if name == "abcmb_los_tangent":
            fl = [w["n_cosmo"] * w["n_ell"] * w["nkt"] * w["n_tau"] * (LOS_T_FLOPS_VALUE + LOS_T_FLOPS_PER_DIR
                                                                       * w["n_dir"]) for _, _, w in calls]
            kern = "los_tangent_kernel (+ its partial-sum kernel)"
        else:
            fl = [w["n_cosmo"] * w["n_mu"] * w["n_l"] * LENS_FLOPS_PER_MU_L for _, _, w in calls]
            kern = ("cgl_kernel + ksi_kernel + back_kernel + back_final_kernel" if calls[0][2]["mode"] == "vector"
                    else "Cgl and back-transform GEMMs on MFMA (FP64) + ksi_kernel")
        a = sum(fl) / (sum(cms) * 1e-3) / 1e12
        others.append({"call": name, "kernel": kern, "bound": "fp64-valu", "calls": len(cms),
                       "ms_per_call": sum(cms) / len(cms), "achieved": a, "peak": FP64_PEAK_TFLOPS,
                       "unit": "TFLOP/s", "frac": a / FP64_PEAK_TFLOPS})
    roof["other_kernels"] = others

    # batched forward spectra (64 Latin-hypercube cosmologies with their own k grids, lensed to l = 2500,
    # FP64 ODE + FP32 LOS): host preparation + one ODE launch + downstream
    from abcmb_amd.model import Model, MassiveNeutrino
    mf = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096, los_precision="fp32")
    if tuned:
        autotune.apply(mf, dict(tuned, lens_mode="vector"))
    nf = 64
    tf_thetas = LK.latin_hypercube(nf, seed=2)
    tf_k = LK.uneven_k_grids(nf, seed=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plans_f = mf.prepare_many([mf.derived(LK.theta_to_params(t)) for t in tf_thetas], k_list=tf_k)
    mf.execute(plans_f)
    torch.cuda.synchronize()
    spectra = nf / (time.perf_counter() - t0)
    print(f"# forward {spectra:.2f} spectra/s", file=sys.stderr, flush=True)

    cpu, cl_err = None, None
    if not args.no_cpu_baseline:
        plan0 = model.prepare(model.derived(LK.theta_to_params(thetas[0])), k=k_grids[0])
        g0 = model.execute([plan0])[0]
        gl = [x.cpu().numpy() for x in (g0.ClTT, g0.ClTE, g0.ClEE, g0.ClBB)]
        cpu, cl_err = cpu_baseline(model, plan0, gl)
        print(f"# C_l error vs oracle {cl_err}", file=sys.stderr, flush=True)
    gates = {"grad_err_vs_converged<=1e-3": grad_err_conv <= 1e-3,
             f"speedup_vs_plain>=10": value / plain >= 10.0}
    if cl_err is not None:
        gates["cl_err_vs_oracle TT/TE/EE<=1e-4"] = max(cl_err["TT"], cl_err["TE"], cl_err["EE"]) <= 1e-4
        gates["cl_err_vs_oracle BB<=1e-3"] = cl_err["BB"] <= 1e-3
    gates["cl_diff_vs_plain<=1e-4"] = cl_diff_plain <= 1e-4
    tol_used = dict(TOL_ACCURATE) if tol else {}

    out = {
        "metric": METRIC, "value": value, "unit": "grad-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched lensed-Cl likelihood gradients, {args.batch} Latin-hypercube cosmologies "
                               f"(massive nu, uneven 300-800 k-grids, l<=2500, FP64 ODE + FP64 LOS)",
                   "ode_tolerances": ("reference's (rtol 1e-5 / 1e-4, atol 1e-10 / 1e-6 below / above k=0.01)"
                                      if not tol else f"reference's with the large-k ones tightened to {tol_used}"),
                   "gradient": (f"tangent-linear ODE (5 directions, one launch per chunk: {args.chunk // 8}, "
                                f"{args.chunk // 4}, {args.chunk // 2}, then {args.chunk} cosmologies) + "
                                + ("downstream differenced along 7 directions" if args.fd_downstream
                                   else "tangent-linear LOS, lensing differenced along 7 directions")
                                if args.gradient == "tangent" else
                                "central finite differences, 15 evaluations, grids frozen, all in one ODE launch"),
                   "autotuned": tuned,
                   "plain_path": "one cosmology per launch, FP64, central finite differences (15 evaluations), "
                                 "the reference's ODE tolerances",
                   "plain_path_grad_evals_per_s": plain, "speedup_vs_plain": value / plain,
                   "grad_diff_vs_plain_rel_to_norm": agree, "grad_diff_vs_plain_max_rel": agree_rel,
                   "grad_err_vs_converged": grad_err_conv,
                   "plain_path_grad_err_vs_converged": plain_err_conv,
                   "reference_tolerances": {"grad_evals_per_s": ref_rate, "grad_err_vs_converged": ref_err_conv,
                                            "workload": f"the same {args.batch} cosmologies, one step"},
                   "large_batch": large,
                   "adjoint": adjoint,
                   "los_variants": los_variants,
                   "los_variants_note": "the tangent-linear LOS with FP32 arithmetic, or FP32 accumulation of "
                                        "bfloat16-rounded operands, in the timed gradient path; errors against "
                                        "the FP64 LOS (spectra of the plain-path cosmologies, gradients of the "
                                        "whole batch relative to each norm) and against the converged gradients",
                   "grad_err_reference": "tangent-linear gradients of the first plain-path cosmologies with the "
                                         "large-k ODE tolerances / 300 and the small-k ones / 30 (max error over "
                                         "parameters relative to the gradient norm, each component in units of "
                                         "its parameter)",
                   "cl_err_vs_oracle": cl_err,
                   "cl_err_reference": "lensed C_l (l=2..2500) of cosmology 0 on the GPU path vs the complete "
                                       "FP64 oracle evaluation of the same inputs; TT/EE/BB relative per l, TE "
                                       "relative to max|TE|",
                   "cl_diff_vs_plain": cl_diff_plain, "L_diff_vs_plain": float(np.max(np.abs(L[:np_] - Lp)
                                                                                       / np.abs(Lp))),
                   "gates": gates,
                   "forward_spectra_per_s": spectra,
                   "forward_workload": "64 Latin-hypercube cosmologies (seed 2), own 300-800 k grids, lensed "
                                       "TT/TE/EE/BB to l=2500, FP64 ODE + FP32 LOS, host preparation included",
                   "phases_s": phases},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
