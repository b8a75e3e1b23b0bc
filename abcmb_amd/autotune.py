"""Autotuner of the launch shapes: picks, by timing on the caller's own
workload, the ODE launch schedule (one workgroup per work item, or a
persistent grid pulling items from a device work queue, at 100% or 200% of
the resident workgroups) and the lensing path (vector units, or the MFMA GEMMs
with 32- or 64-cosmology column tiles), and keeps the choice in an on-disk
cache keyed by the device and the workload's shape.

The settings are process-wide (``abcmb_tune_set`` in the C ABI) plus the
model's ``lens_mode``; ``apply`` sets both.  The reference has no counterpart
(XLA picks its own launch shapes).
"""
from __future__ import annotations

import json
import os
import time

from . import abi

PERT_CANDIDATES = (0, 100, 200)
LENS_CANDIDATES = (("vector", 0), ("mfma", 32), ("mfma", 64))


def cache_path() -> str:
    return os.environ.get("ABCMB_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "abcmb_amd", "autotune.json"))


def _load(path):
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + f".tmp{os.getpid()}"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def workload_key(model, n_cosmo: int, tangent: bool) -> str:
    import torch
    lay = model.layout
    return json.dumps({"device": torch.cuda.get_device_name(0), "n": lay.n, "tangent": bool(tangent),
                       "lensing": bool(model.specs["lensing"]), "n_l": int(model.lensing_ells.size),
                       "batch": int(2 ** round(__import__("math").log2(max(1, n_cosmo))))}, sort_keys=True)


def apply(model, choice: dict):
    lib = abi.hip_lib()
    abi.check(lib.abcmb_tune_set(abi.TUNE_PERT_PERSISTENT, int(choice["pert_persistent"])), "abcmb_tune_set")
    abi.check(lib.abcmb_tune_set(abi.TUNE_LENS_GEMM_COLS, int(choice["lens_gemm_cols"])), "abcmb_tune_set")
    if model.specs["lensing"]:
        model.lens_mode = choice["lens_mode"]


def _time(fn, reps=2):
    """Mean wall time of ``fn`` (GPU work included) over ``reps`` runs after one warm-up."""
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune(model, plans, seeds=None, force: bool = False, path: str = None) -> dict:
    """Choose and apply the launch shapes for ``model`` on a sample of the
    workload (``plans``; ``seeds`` selects the tangent-linear ODE kernel).
    Returns the choice with the measured times; a cached choice for the same
    device and workload shape is applied without timing unless ``force``."""
    path = path or cache_path()
    data = _load(path)
    key = workload_key(model, len(plans), seeds is not None)
    if key in data and not force:
        apply(model, data[key])
        return data[key]
    lib = abi.hip_lib()
    saved_mode = model.lens_mode
    times = {"pert": {}, "lens": {}}
    for sched in PERT_CANDIDATES:
        abi.check(lib.abcmb_tune_set(abi.TUNE_PERT_PERSISTENT, sched), "abcmb_tune_set")
        times["pert"][sched] = _time(lambda: model._run_ode(plans, tangents=seeds), reps=1)
    best_pert = min(times["pert"], key=times["pert"].get)
    abi.check(lib.abcmb_tune_set(abi.TUNE_PERT_PERSISTENT, best_pert), "abcmb_tune_set")
    best_lens = ("vector", 0)
    if model.specs["lensing"] and model.lens_precision == 0:
        run = model._run_ode(plans)
        ds = model._downstream(plans, run["pts"], run["kdev"])
        for mode, cols in LENS_CANDIDATES:
            model.lens_mode = mode
            abi.check(lib.abcmb_tune_set(abi.TUNE_LENS_GEMM_COLS, cols), "abcmb_tune_set")
            times["lens"][f"{mode}/{cols}"] = _time(
                lambda: model.lensed_from_cl(plans, run["pts"], run["kdev"], ds["cl_full"]))
        best_lens = tuple(min(LENS_CANDIDATES, key=lambda c: times["lens"][f"{c[0]}/{c[1]}"]))
    else:
        model.lens_mode = saved_mode
    choice = {"pert_persistent": best_pert, "lens_mode": best_lens[0], "lens_gemm_cols": best_lens[1],
              "times_s": {g: {str(k): v for k, v in t.items()} for g, t in times.items()}}
    apply(model, choice)
    data[key] = choice
    _save(path, data)
    return choice
