"""Benchmark: gradient evaluations per second of the 7-parameter lensed-C_l
likelihood (l <= 2500) on one MI355X.

Workload (batched gradients): ``--batch`` Latin-hypercube cosmologies
(omega_b, omega_cdm, h, ln 1e10 A_s, n_s, tau_reio, sum m_nu over the stated
ranges, seed 0), each with its own uneven log-k grid of 300-800 modes in
[7.05e-6, 0.62] Mpc^-1 (seed 0), one massive neutrino species (3 momentum
bins x l <= 17, state size 100), lensed TT/TE/EE to l = 2500 in FP64 (ODE
and LOS), Gaussian log-likelihood against a per-cosmology seeded mock.
One step = the gradients of all ``--batch`` cosmologies: tangent-linear ODE
launches of ``--chunk`` cosmologies each (perturbation tables and their
derivatives along omega_b, omega_cdm, h, tau_reio, m_nu), each followed by
a tangent-linear LOS pass and the lensing differenced along all seven
directions (``--fd-downstream``: the whole downstream differenced), with the
next chunk's host preparation overlapping the GPU work (``--gradient fd`` instead:
central finite differences, 15 full model evaluations per gradient, all in
one ODE launch).

Also reported: the plain path (one cosmology per launch, FP64, central
finite differences) on a few of the same cosmologies and its agreement with
the batched gradients, the roofline of the dominant kernel (the perturbation
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


def build_model():
    from abcmb_amd.model import Model, MassiveNeutrino
    # max_steps_PE raised from the reference default 2048: with massive neutrinos the
    # highest modes (k ~ 0.6 Mpc^-1) need up to ~2000 steps in this integrator
    return Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096)


def cpu_baseline(model, thetas, k_grids, seconds_budget=20.0):
    """Oracle (CPU FP64 restatement) on a bounded sample of one gradient evaluation."""
    from abcmb_amd.likelihood import theta_to_params
    from oracle import oracle as O
    from oracle import pipeline
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p = model.derived(theta_to_params(thetas[0]))
    plan = model.prepare(p, k=k_grids[0])
    nk = plan.k.size
    sub = np.linspace(0, nk - 1, 16).astype(int)
    t0 = time.perf_counter()
    pipeline.run(model, plan, k_subset=sub)
    t_ode_sub = time.perf_counter() - t0
    t_ode = t_ode_sub * nk / sub.size                   # all k-modes of one cosmology
    # downstream (sources, LOS, lensing) once, with the ODE table from the GPU run
    g = model.execute([plan])[0]
    pt = g.PT.table.cpu().numpy()
    t0 = time.perf_counter()
    src = O.sources(pt, plan.k, model.n_save, plan.bgv, plan.lna_save, model.k_transfer, model.scales)
    cl_raw = O.los_cl(src, plan.bgv[0], plan.bg.tau0, model.k_transfer, model.bes_tab, model.bes_rng, model.bes_l,
                      model.ell_idx, p["A_s"], p["n_s"], model.specs["k_pivot"])
    cl_full = cl_raw @ model.ell_op.T
    c = plan.bg.cosmo
    import abcmb_amd.abi as abi
    cc = abi.AbcmbCosmo.from_buffer_copy(c)
    lna = np.ascontiguousarray(plan.lna_save)
    cc.lna_save, cc.k, cc.pt, cc.n_save, cc.nk = lna.ctypes.data, plan.k.ctypes.data, pt.ctypes.data, model.n_save, nk
    clpp = O.clpp(cc, pt, plan.lens_lna, plan.lens_aH, plan.lens_tau, plan.chi_rec, model.lensing_ells.astype(float))
    O.lens(cl_full, clpp, model.mu, model.w)
    t_down = time.perf_counter() - t0
    t_forward = t_ode + t_down
    t_grad = 15 * t_forward
    return {"value": 1.0 / t_grad, "unit": "grad-evals/s", "cores": threads, "kind": "port",
            "sample": (f"oracle ODE on {sub.size} of {nk} k-modes ({t_ode_sub:.1f} s, scaled by k count) + full "
                       f"sources/LOS/lensing ({t_down:.1f} s) of one cosmology; x15 forward evaluations per "
                       f"central-difference gradient")}


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
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--plain", type=int, default=2, help="cosmologies timed on the plain path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chunk", type=int, default=64, help="cosmologies per tangent-linear ODE launch")
    ap.add_argument("--gradient", choices=("tangent", "fd"), default="tangent",
                    help="batched gradient method (the plain path is always finite differences)")
    ap.add_argument("--fd-downstream", action="store_true",
                    help="difference the whole downstream along each direction instead of differentiating the "
                         "LOS in one tangent-linear pass")
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
    from abcmb_amd import likelihood as LK

    t_setup = time.perf_counter()
    model = build_model()
    thetas = LK.latin_hypercube(args.batch, seed=0)
    k_grids = LK.uneven_k_grids(args.batch, seed=0)
    # mock: C_l at the fiducial point, noise per cosmology slot
    fid = model.execute([model.prepare(model.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(model, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=args.batch, seed=1)
    print(f"# setup {time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)

    phases = {}
    if args.gradient == "tangent":
        batched = lambda: LK.tangent_gradients(model, like, thetas, k_grids=k_grids, phases=phases, chunk=args.chunk,
                                               downstream="fd" if args.fd_downstream else "tangent")
    else:
        batched = lambda: LK.fd_gradients(model, like, thetas, k_grids=k_grids, batched=True)
    for w in range(args.warmup):
        batched()
        print(f"# warmup {w} done", file=sys.stderr, flush=True)

    model.timing = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        L, grad = batched()
        print(f"# step {s} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timing = model.timing
    model.timing = None
    value = args.batch * args.steps / dt

    # plain path: one cosmology per launch, same FD gradient
    t0 = time.perf_counter()
    Lp, gp = LK.fd_gradients(model, like, thetas[:args.plain], k_grids=k_grids[:args.plain], batched=False)
    torch.cuda.synchronize()
    plain = args.plain / (time.perf_counter() - t0)
    gb = grad[:args.plain]
    # agreement of the batched gradient with the plain path's, per parameter, relative to the
    # gradient's norm in units of each parameter's magnitude (components near zero stay meaningful)
    scale = np.linalg.norm(gp * thetas[:args.plain], axis=1, keepdims=True) / np.abs(thetas[:args.plain])
    agree = float(np.max(np.abs(gp - gb) / scale))
    agree_rel = float(np.max(np.abs(gp - gb) / (np.abs(gp) + 1e-300)))
    print(f"# plain path {plain:.4f} grad-evals/s", file=sys.stderr, flush=True)

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
        traffic = t["fetch_bytes_per_launch"] + t["write_bytes_per_launch"]
    roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic, "kernel": kname,
            "launch_ms": ms, "flops_per_launch": flops,
            "chord_iters_per_step": sum(t["iters"] for t in timing) / max(1, sum(t["steps"] for t in timing)),
            "traffic_source": "profiles/pert_tangent_traffic.json (FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                              "scaled by a 1 GiB copy)" if traffic else None}

    # batched forward spectra (64 Latin-hypercube cosmologies with their own k grids, lensed to l = 2500,
    # FP64 ODE + FP32 LOS): host preparation + one ODE launch + downstream
    from abcmb_amd.model import Model, MassiveNeutrino
    mf = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096, los_precision="fp32")
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

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(model, thetas, k_grids)

    out = {
        "metric": METRIC, "value": value, "unit": "grad-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"batched lensed-Cl likelihood gradients, {args.batch} Latin-hypercube cosmologies "
                               f"(massive nu, uneven 300-800 k-grids, l<=2500, FP64 ODE + FP64 LOS)",
                   "gradient": (f"tangent-linear ODE (5 directions, one launch per chunk: {args.chunk // 8}, "
                                f"{args.chunk // 4}, {args.chunk // 2}, then {args.chunk} cosmologies) + "
                                + ("downstream differenced along 7 directions" if args.fd_downstream
                                   else "tangent-linear LOS, lensing differenced along 7 directions")
                                if args.gradient == "tangent" else
                                "central finite differences, 15 evaluations, grids frozen, all in one ODE launch"),
                   "plain_path": "one cosmology per launch, FP64, central finite differences (15 evaluations)",
                   "plain_path_grad_evals_per_s": plain, "speedup_vs_plain": value / plain,
                   "grad_diff_vs_plain_rel_to_norm": agree, "grad_diff_vs_plain_max_rel": agree_rel,
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
