"""``Model`` / ``Output`` / ``PerturbationTable``: the drop-in API.

Mirrors the reference's user-facing surface (abcmb/main.py:29-184,
:614-650; abcmb/perturbations.py:406-450): ``Model(user_species, **specs)``
then ``model(params) -> Output`` with the same field names and units
(C_l dimensionless, P(k) in Mpc^3, k in Mpc^-1).  Background and
recombination tables come from the host producer; the perturbation ODE,
sources, line-of-sight transfer, C_l quadrature, l-spline, lensing and
P(k) run in ``libabcmb_hip.so`` on the GPU.  ``Model.run_batch`` runs many
cosmologies with one ODE launch for all their k-modes.

Host-side preparation (``Model.prepare``) and device execution
(``Model.execute``) are separate so that every cosmology's inputs are fixed
before any kernel runs.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import abi, bessel, splines
from . import species as SPC
from . import specs as S
from .background import Background
from .params import add_derived_parameters

LCDM_SPECIES = tuple(c.name for c in SPC.LCDM_SPECIES)
MassiveNeutrino = SPC.MassiveNeutrino


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("abcmb_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


@dataclass
class CosmoPlan:
    """Everything the GPU path needs for one cosmology, prepared on the host."""
    params: dict
    bg: Background
    lna_save: np.ndarray
    lna_start: np.ndarray
    rtol: np.ndarray
    atol: np.ndarray
    bgv: np.ndarray               # [6][n_save-1] LOS background vectors
    lens_lna: np.ndarray = None   # [500] Limber ln a axis
    lens_aH: np.ndarray = None
    lens_tau: np.ndarray = None
    chi_rec: float = 0.0
    k: np.ndarray = None          # PT k-grid of this cosmology (Mpc^-1)


@dataclass
class TangentSeeds:
    """Derivatives of one cosmology's host inputs along the tangent directions.

    par: [N_TANGENT_DIRS][COSMO_DOUBLES] derivatives of the AbcmbCosmo double fields;
    tau, xe, tm: [N_TANGENT_DIRS][len] derivatives of the tables.  Directions
    beyond the ones a caller asks for are zero.
    """
    par: np.ndarray
    tau: np.ndarray
    xe: np.ndarray
    tm: np.ndarray


_TAU0_FIELD = [f for f, _ in abi.AbcmbCosmo._fields_].index("tau0")   # its index among the double fields


def cosmo_doubles(c) -> np.ndarray:
    """The COSMO_DOUBLES double fields of an AbcmbCosmo, in declaration order."""
    return np.frombuffer(bytes(c), dtype=np.float64, count=abi.COSMO_DOUBLES).copy()


def tangent_seeds(plus, minus, steps) -> TangentSeeds:
    """Central differences of the host producer between plans at theta +- step.

    ``plus[d]``/``minus[d]`` were prepared with every grid frozen at the base
    plan, so the tables share their abscissae and the differences are smooth
    functions of the parameters.  The producer (background integrals,
    recombination, reionization root) is deterministic and converged to
    rounding, so a relative step of 1e-4 leaves differences accurate to ~1e-8.
    """
    nd = abi.N_TANGENT_DIRS
    if not (len(plus) == len(minus) == len(steps) <= nd):
        raise ValueError(f"need at most {nd} directions with a plus and a minus plan each")
    p0 = plus[0].bg
    par = np.zeros((nd, abi.COSMO_DOUBLES))
    tau = np.zeros((nd, p0.tau_tab.size))
    xe = np.zeros((nd, p0.xe_tab.size))
    tm = np.zeros((nd, p0.tm_tab.size))
    for d, (pp, pm, h) in enumerate(zip(plus, minus, steps)):
        if pp.bg.xe_tab.size != xe.shape[1] or pm.bg.xe_tab.size != xe.shape[1]:
            raise ValueError("plus/minus plans have different recombination grids")
        s = 1.0 / (2.0 * h)
        par[d] = (cosmo_doubles(pp.bg.cosmo) - cosmo_doubles(pm.bg.cosmo)) * s
        tau[d] = (pp.bg.tau_tab - pm.bg.tau_tab) * s
        xe[d] = (pp.bg.xe_tab - pm.bg.xe_tab) * s
        tm[d] = (pp.bg.tm_tab - pm.bg.tm_tab) * s
    return TangentSeeds(par, tau, xe, tm)


class _Modes:
    """``modes`` of a species' output_perturbations (abcmb/perturbations.py:318-404):
    modes[i] is state row i on (Nlna, Nk) for the user-fluid rows (kept by the
    kernels as they are) and the standard rows the table stores."""

    def __init__(self, pt_fluid, first_row, std):
        self.pf, self.first, self.std = pt_fluid, first_row, std

    def __getitem__(self, i):
        i = int(i)
        if self.pf is not None and self.first <= i < self.first + self.pf.shape[0]:
            return self.pf[i - self.first]
        if i in self.std:
            return self.std[i]
        raise KeyError(f"state row {i} is not kept in the perturbation table")


class PerturbationTable:
    """Perturbation table (abcmb/perturbations.py:406-450); arrays are (Nlna, Nk) views."""

    def __init__(self, k, lna, pt, has_ncdm, species_off=0, fluids=None, pt_fluid=None, args=None):
        self.k = k
        self.lna = lna
        self.table = pt  # [21][Nlna][Nk]
        I = abi.PT_INDEX
        self.delta_m = pt[I["delta_m"]]
        self.theta_b_prime = pt[I["theta_b_prime"]]
        self.metric_eta = pt[I["metric_eta"]]
        self.metric_h_prime = pt[I["metric_h_prime"]]
        self.metric_eta_prime = pt[I["metric_eta_prime"]]
        self.metric_alpha = pt[I["metric_alpha"]]
        self.metric_alpha_prime = pt[I["metric_alpha_prime"]]
        sp = {
            "DarkEnergy": {},
            "ColdDarkMatter": {"delta": pt[I["cdm_delta"]]},
            "Baryon": {"delta": pt[I["b_delta"]], "theta": pt[I["b_theta"]]},
            "Photon": {"delta": pt[I["g_delta"]], "theta": pt[I["g_theta"]], "sigma": pt[I["g_sigma"]],
                       "G0": pt[I["g_G0"]], "G2": pt[I["g_G2"]]},
            "MasslessNeutrino": {"delta": pt[I["ur_delta"]], "theta": pt[I["ur_theta"]],
                                 "sigma": pt[I["ur_sigma"]]},
        }
        if has_ncdm:
            sp["MassiveNeutrino"] = {"delta": pt[I["ncdm_delta"]], "theta": pt[I["ncdm_theta"]],
                                     "sigma": pt[I["ncdm_sigma"]]}
        if species_off & abi.NO_LAMBDA:
            del sp["DarkEnergy"]
        if species_off & abi.NO_CDM:
            del sp["ColdDarkMatter"]
        if fluids is not None:
            modes = _Modes(pt_fluid, fluids.first_row, {1: pt[I["cdm_delta"]], 2: pt[I["b_delta"]],
                                                        3: pt[I["b_theta"]]})
            for f in fluids.user:
                sp[f.name] = f.output_perturbations(lna, modes, args)
        self.pt_fluid = pt_fluid
        self.species_perturbations = sp


@dataclass
class Output:
    """Result of one cosmology (abcmb/main.py:614-650) plus ClBB and C_l^phiphi."""
    ClTT: object
    ClTE: object
    ClEE: object
    Pk: object
    l: np.ndarray
    k: np.ndarray
    BG: Background
    PT: PerturbationTable
    params: dict
    ClBB: object = None
    Clpp: object = None
    steps: object = None


class Model:
    """GPU drop-in for ``abcmb.Model`` on the perturbation + LOS path."""

    def __init__(self, user_species=None, los_precision: str = "fp64", lens_mode: str = "vector",
                 lens_precision: str = "fp64", device_background: bool = False, **kwargs):
        kwargs.pop("adjoint", None)  # reference option for diffrax AD; not used here
        specs = S.load_specs(kwargs)
        self.specs = specs
        self.fixed_params = {}
        self._species_setup(user_species)
        k_pt, k_pk = S.get_k_axis_perturbations(specs)
        if specs.get("k_axis_perturbations") is not None:
            k_pt = np.asarray(specs["k_axis_perturbations"], dtype=np.float64)
        self.k_pt = np.ascontiguousarray(k_pt)
        self.k_pk = np.ascontiguousarray(k_pk)
        self.k_transfer = np.ascontiguousarray(S.get_k_axis_transfer(specs))
        self.n_save = int(specs["n_lna_save"])
        if specs["l_min"] != 2:
            raise NotImplementedError("l_min must be 2")
        # multipoles (abcmb/spectrum.py:180-233)
        self.bes_l, self.bes_tab, self.bes_rng = bessel.load_tables()
        l_min, l_max = specs["l_min"], specs["l_max"]
        i_lo = int(np.where(self.bes_l <= l_min)[0][-1])
        self.ells = np.arange(l_min, l_max + 1)
        if specs["lensing"]:
            lens_lmax = l_max + 500
            i_hi = int(np.where(self.bes_l >= lens_lmax)[0][0])
            self.lensing_ells = np.arange(l_min, lens_lmax + 1)
            mu, w = np.polynomial.legendre.leggauss(lens_lmax + 70)
            self.mu = np.concatenate((mu, [1.0]))
            self.w = np.concatenate((w, [0.0]))
        else:
            i_hi = int(np.where(self.bes_l >= l_max)[0][0])
            self.lensing_ells = self.ells
            self.mu = self.w = None
        self.ell_idx = np.arange(i_lo, i_hi + 1, dtype=np.int32)
        self.ell_sampled = self.bes_l[self.ell_idx].astype(np.float64)
        self.ell_op = splines.operator(self.ell_sampled, self.lensing_ells.astype(np.float64))
        self.scales = np.array([specs["scale_sw"], specs["scale_isw"], specs["scale_dop"], specs["scale_pol"]],
                               dtype=np.float64)
        lay = self.layout
        self.settings = abi.AbcmbOdeSettings(
            pcoeff=specs["pcoeff_PE"], icoeff=specs["icoeff_PE"], dcoeff=specs["dcoeff_PE"], dt0=1e-2,
            safety=0.9, factormin=0.2, factormax=10.0, newton_kappa=1e-2, jac_offset=0.6,
            max_steps=int(specs["max_steps_PE"]), max_newton=10,
            l_g=lay.l_g, l_pol=lay.l_pol, l_ur=lay.l_ur, l_ncdm=lay.l_ncdm, n_q=lay.n_q,
            n=lay.n + self.n_fluid, n_fluid=self.n_fluid)
        precisions = {"fp64": 0, "fp32": 1, "bf16": 2}
        if los_precision not in precisions:
            raise ValueError("los_precision must be 'fp64', 'fp32' or 'bf16'")
        self.los_precision = precisions[los_precision]
        # lensing: "vector" (Wigner recursions in registers, FP64) or "mfma" (the Cgl
        # and back-transform contractions as GEMMs on the matrix cores against
        # d-matrices built once, in lens_precision)
        if lens_mode not in ("vector", "mfma") or lens_precision not in precisions:
            raise ValueError("lens_mode must be 'vector' or 'mfma', lens_precision 'fp64', 'fp32' or 'bf16'")
        if lens_mode == "vector" and lens_precision != "fp64":
            raise ValueError("the vector lensing path is FP64 only")
        self.lens_mode = lens_mode
        self.lens_precision = precisions[lens_precision]
        # device_background: the host producer stops after recombination; the
        # reionization root (abcmb_reionize) and the LOS background vectors with the
        # optical depth (abcmb_los_background) run on the GPU before the ODE launch
        self.device_background = bool(device_background)
        # host_tangents: how the tangent seeds of the host inputs are made --
        # "forward": the producer's forward-mode derivatives (tangent_plans);
        # "fd": central differences of producer runs with the grids held
        self.host_tangents = "forward"
        self.chunk = 128   # cosmologies per downstream launch (bounds the source buffers)
        self.timing = None  # list: per-launch pert_fwd event timing and work counts when set
        self.call_timing = None  # dict: entry-point name -> [(start event, end event, work)] when set
        self._dev = None

    def _species_setup(self, user_species):
        """The model's species (abcmb/model_specs.py:84-118): the LCDM set unless
        use_LCDM_species is off, then the user's classes, instantiated in that
        order.  The standard species' rows are the kernels' fixed layout; user
        fluids (abcmb_amd.species) take the rows after it and are compiled into
        a build of the libraries for their set (abcmb_amd.fluidgen)."""
        specs = self.specs
        classes = list(SPC.LCDM_SPECIES) if specs.get("use_LCDM_species", True) else []
        for s in (user_species or ()):
            if isinstance(s, str) and s == "MassiveNeutrino":
                s = SPC.MassiveNeutrino
            if not isinstance(s, type):
                raise TypeError(f"user_species takes Fluid classes, not {s!r} (the model instantiates them)")
            if not issubclass(s, SPC.Fluid):
                raise TypeError(f"{s.__name__} is not a subclass of abcmb_amd.species.Fluid")
            classes.append(s)
        std = {c.name for c in classes if SPC.is_standard(c)}
        missing = [nm for nm in ("Baryon", "Photon", "MasslessNeutrino") if nm not in std]
        if missing:
            raise NotImplementedError(f"the kernels need the standard species {missing} in the model")
        self.has_ncdm = "MassiveNeutrino" in std
        self.layout = lay = S.state_layout(specs, self.has_ncdm)
        rows = {"DarkEnergy": 1, "ColdDarkMatter": lay.i_cdm, "Baryon": lay.i_b, "Photon": lay.i_g,
                "MasslessNeutrino": lay.i_ur, "MassiveNeutrino": lay.i_ncdm}
        objs, nxt = [], lay.n
        for c in classes:
            if SPC.is_standard(c):
                objs.append(c(rows[c.name], specs))
            else:
                o = c(nxt, specs)
                o.first_idx = nxt
                nxt += int(o.num_equations)
                objs.append(o)
        names = [o.name for o in objs]
        if len(set(names)) != len(names):
            raise ValueError(f"species names must be distinct: {names}")
        self.species_objs = tuple(objs)
        self.species = tuple(o.name if SPC.is_standard(o) else o for o in objs)
        self.species_off = (0 if "DarkEnergy" in std else abi.NO_LAMBDA) | (0 if "ColdDarkMatter" in std
                                                                           else abi.NO_CDM)
        self.n_fluid = nxt - lay.n
        user = [o for o in objs if not SPC.is_standard(o)]
        self.fluids = None
        self._hip_path = self._host_path = None
        if user:
            from .fluidgen import FluidSet
            self.fluids = FluidSet(objs, user, lay.n, nxt, self.species_off)
            self.fluids.build(hip=False)   # the host producer (seconds); the kernels on first GPU use
            self._hip_path, self._host_path = self.fluids.hip_path, self.fluids.host_path
            for lib, fn in ((self._hlib(), "abcmb_host_fluid_signature"), ):
                sig = getattr(lib, fn)().decode()
                if sig != self.fluids.signature:
                    raise RuntimeError(f"{self._host_path} was built for fluids {sig!r}, "
                                       f"not {self.fluids.signature!r}")

    def _lib(self):
        """The kernel library of this model's species (a fluid build when it has user fluids,
        compiled here the first time: the ODE kernels take a few minutes)."""
        if self.fluids is not None and not self.fluids.built():
            if not self.specs.get("build_fluids", True):
                raise abi.HipLibraryMissing(f"the kernels for fluids {[f.name for f in self.fluids.user]} are not "
                                            f"built ({self.fluids.directory})")
            print(f"building the kernels for user fluids {[f.name for f in self.fluids.user]} "
                  f"({self.fluids.signature}); a few minutes, once", flush=True)
            self.fluids.build()
        lib = abi.hip_lib(self._hip_path)
        if self.fluids is not None and lib.abcmb_fluid_signature().decode() != self.fluids.signature:
            raise RuntimeError(f"{self._hip_path} was built for other fluids")
        return lib

    def _hlib(self):
        return abi.host_lib(self._host_path)

    def fix_params(self, **params):
        """Parameters merged into every parameter set this model derives (a call's
        own keys win): the user fluids' parameters when the caller varies only the
        standard ones (the likelihood and gradient helpers)."""
        self.fixed_params.update(params)
        return self

    # ------------------------------------------------------------------ host side
    def derived(self, params: dict) -> dict:
        return add_derived_parameters({**self.fixed_params, **params}, self.species, self.specs)

    def prepare(self, full_params: dict, freeze: CosmoPlan = None, k: np.ndarray = None) -> CosmoPlan:
        """Host background and ODE inputs of one cosmology.

        ``freeze``: reuse that plan's discrete grid choices (save grid from
        lna_transfer_start, lna_rec, per-k start times, recombination switch)
        so that finite differences see a smooth function of the parameters.
        ``k``: PT k-grid of this cosmology (default: the model's).
        """
        sp = self.specs
        kk = np.ascontiguousarray(self.k_pt if k is None else k, dtype=np.float64)
        bg = Background(full_params, self.has_ncdm, n_q=self.layout.n_q, k_pivot=sp["k_pivot"],
                        input_tau_reion=sp["input_tau_reion"], species_off=self.species_off,
                        fluid_par=self.fluids.fluid_par(full_params) if self.fluids is not None else None,
                        lib=self._hlib(),
                        rec_switch=freeze.bg.rec_switch if freeze is not None else None,
                        lna_rec=freeze.bg.lna_rec if freeze is not None else None,
                        lna_transfer_start=freeze.bg.lna_transfer_start if freeze is not None else None,
                        reionization=not self.device_background)
        if freeze is not None:
            bg.cosmo.lna_rec = freeze.bg.lna_rec
            bg.cosmo.lna_transfer_start = freeze.bg.lna_transfer_start
            lna_save = freeze.lna_save
            ls = freeze.lna_start
        else:
            lna_save = np.linspace(bg.lna_transfer_start, 0.0, self.n_save)
            ls = bg.lna_start(kk, sp["R_tc"], sp["R_large"])
        big = kk > sp["k_split_PE"]
        rtol = np.where(big, sp["rtol_large_k_PE"], sp["rtol_small_k_PE"])
        atol = np.where(big, sp["atol_large_k_PE"], sp["atol_small_k_PE"])
        bgv = None if self.device_background else bg.los_vectors(lna_save[:-1])
        plan = CosmoPlan(full_params, bg, lna_save, ls, rtol, atol, bgv, k=kk)
        if sp["lensing"]:
            lens_lna = np.linspace(bg.lna_rec, 0.0, 500)
            ev = bg.eval(lens_lna)
            plan.lens_lna = lens_lna
            plan.lens_aH = ev["aH"]
            plan.lens_tau = ev["tau"]
            plan.chi_rec = bg.tau0 - float(bg.eval(np.array([bg.lna_rec]))["tau"][0])
        return plan

    def tangent_plans(self, base: CosmoPlan, pm_params, steps) -> list:
        """Plans at theta +- step along each direction, from the host producer's
        forward-mode derivatives at ``base`` (every grid choice held): every
        table, double field, LOS vector and lensing background of plan (d, +-)
        is base's +- step[d] times its exact derivative, and its parameters are
        ``pm_params`` (the derived parameters at theta +- step, in the order
        (d, +), (d, -)).  Differences of these plans are the producer's
        derivatives to rounding; the derived parameters themselves (closed forms
        and the BBN helium table, ``params.add_derived_parameters``) enter by
        their central differences.  Same order and meaning as ``prepare`` with
        ``freeze=base`` at each of ``pm_params``."""
        from .background import input_doubles
        fpar = (lambda q: self.fluids.fluid_par(q)) if self.fluids is not None else (lambda q: None)
        nd = len(steps)
        N = abi.N_TANGENT_DIRS
        if len(pm_params) != 2 * nd or nd > N:
            raise ValueError(f"need a plus and a minus parameter set for each of at most {N} directions")
        kp = self.specs["k_pivot"]
        dpar, dtr, dne = np.zeros((N, abi.COSMO_DOUBLES)), np.zeros(N), np.zeros(N)
        for d in range(nd):
            pp, pm = pm_params[2 * d], pm_params[2 * d + 1]
            s = 1.0 / (2.0 * steps[d])
            dpar[d] = (input_doubles(pp, kp, fpar(pp)) - input_doubles(pm, kp, fpar(pm))) * s
            dtr[d] = (pp.get("tau_reion", 0.0) - pm.get("tau_reion", 0.0)) * s
            dne[d] = (pp["Neff"] - pm["Neff"]) * s
        bg = base.bg
        t = bg.tangents(dpar, dtr, dne, reionization=not self.device_background)
        dbgv = None if base.bgv is None else bg.los_vectors_tangent(t, base.lna_save[:-1])[1]
        lensing = bool(self.specs["lensing"])
        if lensing:
            dev = bg.eval_tangent(t, np.append(base.lens_lna, bg.lna_rec))[1]
        tau0 = _TAU0_FIELD
        out = []
        for d in range(nd):
            for si, sgn in enumerate((1.0, -1.0)):
                h = sgn * steps[d]
                params = pm_params[2 * d + si]
                plan = dataclasses.replace(base, params=params, bg=bg.shifted(params, h, t, d),
                                           bgv=None if dbgv is None else base.bgv + h * dbgv[d])
                if lensing:
                    plan.lens_aH = base.lens_aH + h * dev[d, 0, :-1]
                    plan.lens_tau = base.lens_tau + h * dev[d, 1, :-1]
                    plan.chi_rec = base.chi_rec + h * (t.par[d, tau0] - dev[d, 1, -1])
                out.append(plan)
        return out

    def prepare_many(self, params_list, freeze_list=None, k_list=None, threads: int = None):
        """``prepare`` for many cosmologies on a host thread pool (the producer releases the GIL)."""
        import os
        from concurrent.futures import ThreadPoolExecutor
        n = len(params_list)
        fz = freeze_list or [None] * n
        ks = k_list or [None] * n
        threads = threads or min(16, os.cpu_count() or 1)
        with ThreadPoolExecutor(threads) as ex:
            return list(ex.map(lambda i: self.prepare(params_list[i], fz[i], ks[i]), range(n)))

    # ------------------------------------------------------------------ device side
    def _device_constants(self):
        if self._dev is None:
            torch = _torch()
            lib = self._lib()
            abi.check(lib.abcmb_device_init(), "abcmb_device_init")
            dev = torch.device("cuda")
            t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
            self._dev = dict(
                k_pt=t(self.k_pt), k_pk=t(self.k_pk), kt=t(self.k_transfer),
                bes=t(self.bes_tab), rng=t(self.bes_rng), bes_l=t(self.bes_l, torch.int32),
                ell_idx=t(self.ell_idx, torch.int32), ell_op=t(self.ell_op),
                lens_ells=t(self.lensing_ells.astype(np.float64)),
                mu=t(self.mu) if self.mu is not None else None,
                w=t(self.w) if self.w is not None else None,
                scales=self.scales,
            )
        return self._dev

    def _dmat(self):
        """The lensing d-matrices for the MFMA path in ``lens_precision`` (built on
        first use, then kept)."""
        D = self._device_constants()
        key = ("dmat", self.lens_precision)
        if key not in D:
            torch = _torch()
            lib = self._lib()
            nl = self.lensing_ells.size
            nb = lib.abcmb_lens_dmat_size(self.mu.size, nl, self.lens_precision)
            dm = torch.empty(nb, dtype=torch.uint8, device=torch.device("cuda"))
            abi.check(lib.abcmb_lens_dmat(D["mu"].data_ptr(), self.mu.size, nl, self.lens_precision, dm.data_ptr(), nb,
                                          torch.cuda.current_stream().cuda_stream), "abcmb_lens_dmat")
            D[key] = dm
        return D[key]

    def complete_backgrounds(self, plans, seeds=None, dtarget=None):
        """Finish recombination-only plans (``device_background``) on the GPU.

        abcmb_reionize solves each plan's reionization redshift for its
        tau_reion and adds the reionization to the x_e table; abcmb_los_background
        then evaluates the LOS background vectors (tau, g, g', aH, exp(-kappa),
        aH' aH) on the plan's save grid from the finished table.  Both run on a
        stream of their own, and the results land in the plans' host
        backgrounds (x_e table, z_reion, x_e today) and ``bgv``, as the host
        producer would have left them.  ``seeds`` (one ``TangentSeeds`` per plan,
        of recombination-only tables) and ``dtarget`` ([n][N_TANGENT_DIRS]:
        d tau_reion along each direction) add the implicit-function derivative
        of z_reion and the tangents of the finished table to the seeds.
        Returns the LOS vectors' tangents [n][N_TANGENT_DIRS][6][n_save-1] with
        seeds, else None.
        """
        if not plans:
            return None
        torch = _torch()
        lib = self._lib()
        self._device_constants()
        dev = torch.device("cuda")
        n, nd = len(plans), abi.N_TANGENT_DIRS
        xe_n = plans[0].bg.xe_tab.size
        n_rec = xe_n
        if any(pl.bg.xe_tab.size != xe_n for pl in plans):
            raise ValueError("complete_backgrounds: the plans' x_e tables differ in length")
        nl = self.n_save - 1
        f_xe_last = abi.AbcmbCosmo.xe_last.offset // 8
        f_z = abi.AbcmbCosmo.z_reion.offset // 8
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            stage = _Stage(torch, dev)
            dt = stage.add
            stream = side.cuda_stream
            xe_out = torch.empty((n, xe_n), dtype=torch.float64, device=dev)
            z_out = torch.empty(n, dtype=torch.float64, device=dev)
            status = torch.empty(n, dtype=torch.int32, device=dev)
            bgv = torch.empty((n, 6, nl), dtype=torch.float64, device=dev)
            rec, fin = [], []
            for i, pl in enumerate(plans):
                c = abi.AbcmbCosmo.from_buffer_copy(pl.bg.cosmo)
                c.tau_tab = dt(pl.bg.tau_tab).data_ptr()
                c.xe_tab = dt(pl.bg.xe_tab).data_ptr()
                c.tm_tab = dt(pl.bg.tm_tab).data_ptr()
                rec.append(c)
                f = abi.AbcmbCosmo.from_buffer_copy(c)
                f.xe_tab = xe_out[i].data_ptr()
                fin.append(f)
            rec_dev = self._struct_array(torch, dev, rec, dt)
            fin_dev = self._struct_array(torch, dev, fin, dt)
            reion_dev = self._struct_array(torch, dev, [pl.bg.reion for pl in plans], dt)
            lna = dt(np.stack([pl.lna_save[:-1] for pl in plans]))
            tan_rec = tan_fin = None
            if seeds is not None:
                if dtarget is None or len(seeds) != n:
                    raise ValueError("complete_backgrounds: seeds need one entry per plan and dtarget")
                par = dt(np.stack([sd.par for sd in seeds]))
                par_fin = torch.empty_like(par)
                dxe_out = torch.empty((n, nd, xe_n), dtype=torch.float64, device=dev)
                dz_out = torch.empty((n, nd), dtype=torch.float64, device=dev)
                dbgv = torch.empty((n, nd, 6, nl), dtype=torch.float64, device=dev)
                dtar = dt(np.asarray(dtarget, dtype=np.float64).reshape(n, nd))
                tr, tf = [], []
                for i, sd in enumerate(seeds):
                    tau_s, tm_s = dt(sd.tau), dt(sd.tm)
                    tr.append(abi.AbcmbTangent(par=par[i].data_ptr(), tau_tab=tau_s.data_ptr(),
                                               xe_tab=dt(sd.xe).data_ptr(), tm_tab=tm_s.data_ptr(), pt=None))
                    tf.append(abi.AbcmbTangent(par=par_fin[i].data_ptr(), tau_tab=tau_s.data_ptr(),
                                               xe_tab=dxe_out[i].data_ptr(), tm_tab=tm_s.data_ptr(), pt=None))
                tan_rec = self._struct_array(torch, dev, tr, dt)
                tan_fin = self._struct_array(torch, dev, tf, dt)
            stage.flush()
            abi.check(lib.abcmb_reionize(rec_dev.data_ptr(), tan_rec.data_ptr() if seeds is not None else None, n,
                                         reion_dev.data_ptr(), dtar.data_ptr() if seeds is not None else None, n_rec,
                                         xe_out.data_ptr(), dxe_out.data_ptr() if seeds is not None else None,
                                         z_out.data_ptr(), dz_out.data_ptr() if seeds is not None else None,
                                         status.data_ptr(), stream), "abcmb_reionize")
            # the finished structs take z_reion and x_e today from the device results
            fv = fin_dev.view(torch.float64).view(n, -1)
            fv[:, f_z] = z_out
            fv[:, f_xe_last] = xe_out[:, n_rec - 1]
            if seeds is not None:
                par_fin.copy_(par)
                par_fin[:, :, f_z] = dz_out
                par_fin[:, :, f_xe_last] = dxe_out[:, :, n_rec - 1]
            abi.check(lib.abcmb_los_background(fin_dev.data_ptr(), tan_fin.data_ptr() if seeds is not None else None,
                                               n, lna.data_ptr(), nl, bgv.data_ptr(),
                                               dbgv.data_ptr() if seeds is not None else None, stream),
                      "abcmb_los_background")
            st = status.cpu().numpy()
            if np.any(st != 0):
                raise RuntimeError(f"abcmb_reionize: no reionization redshift in [0, 50] for plans "
                                   f"{np.nonzero(st)[0][:8].tolist()}")
            xe_h, z_h, bgv_h = xe_out.cpu().numpy(), z_out.cpu().numpy(), bgv.cpu().numpy()
            if seeds is not None:
                dxe_h, dz_h, dbgv_h = dxe_out.cpu().numpy(), dz_out.cpu().numpy(), dbgv.cpu().numpy()
        for i, pl in enumerate(plans):
            pl.bg.xe_tab[:] = xe_h[i]
            pl.bg.cosmo.z_reion = float(z_h[i])
            pl.bg.cosmo.xe_last = float(xe_h[i, n_rec - 1])
            pl.bgv = bgv_h[i]
            if seeds is not None:
                sd = seeds[i]
                sd.xe[:] = dxe_h[i]
                sd.par[:, f_z] = dz_h[i]
                sd.par[:, f_xe_last] = dxe_h[i, :, n_rec - 1]
        return dbgv_h if seeds is not None else None

    def _complete_pending(self, plans):
        todo = [pl for pl in plans if pl.bgv is None]
        if todo:
            self.complete_backgrounds(todo)

    def run_batch(self, params_list, derived: bool = False):
        """Outputs for several cosmologies; one ODE launch covers all their k-modes."""
        plist = [p if derived else self.derived(p) for p in params_list]
        return self.execute(self.prepare_many(plist))

    def _cosmo_array(self, plans, pts, kdev, dt, tables=True, pfs=None):
        """AbcmbCosmo structs (device pointers) for ``plans`` with perturbation tables
        ``pts`` (and user-fluid row tables ``pfs``)."""
        ns = self.n_save
        cos = []
        for i, (pl, pt, kd) in enumerate(zip(plans, pts, kdev)):
            bg = pl.bg
            c = abi.AbcmbCosmo.from_buffer_copy(bg.cosmo)
            if tables:
                c.tau_tab = dt(bg.tau_tab).data_ptr()
                c.xe_tab = dt(bg.xe_tab).data_ptr()
                c.tm_tab = dt(bg.tm_tab).data_ptr()
            else:   # downstream kernels read parameters, k, the save grid and the table only
                c.tau_tab = c.xe_tab = c.tm_tab = None
            c.lna_save = dt(pl.lna_save).data_ptr()
            c.k = kd.data_ptr()
            c.n_save, c.nk = ns, int(kd.numel())
            c.pt = pt.data_ptr()
            c.pt_fluid = pfs[i].data_ptr() if pfs is not None else None
            cos.append(c)
        return cos

    @staticmethod
    def _struct_array(torch, dev, structs, dt):
        arr = (type(structs[0]) * len(structs))(*structs)
        return dt(np.frombuffer(bytes(arr), dtype=np.uint8), torch.uint8)

    def _run_ode(self, plans, tangents=None, want_state: bool = False):
        """Perturbation tables of ``plans`` (one ODE launch for all their k-modes).

        ``tangents``: one ``TangentSeeds`` per plan selects the tangent-linear
        kernel, which also returns the tables' derivatives along the seeds'
        directions ([N_TANGENT_DIRS][21][n_save][nk] per cosmology).
        """
        return self._finish_ode(self._launch_ode(plans, tangents, want_state))

    def _launch_ode(self, plans, tangents=None, want_state: bool = False):
        """Enqueue the ODE launch of ``_run_ode`` on the current stream and return
        a handle for ``_finish_ode``; the handle keeps every buffer the launch
        reads or writes alive."""
        self._complete_pending(plans)
        torch = _torch()
        lib = self._lib()
        D = self._device_constants()
        dev = torch.device("cuda")
        stream = torch.cuda.current_stream().cuda_stream
        ns = self.n_save
        stage = _Stage(torch, dev)
        keep = [stage]
        dt = stage.add

        allk = [(pl.k if pl.k is not None else self.k_pt) for pl in plans]
        kdev = [D["k_pt"] if kk is self.k_pt else dt(kk) for kk in allk]
        pts = [torch.empty((abi.N_PT_FIELDS, ns, kk.size), dtype=torch.float64, device=dev) for kk in allk]
        nf = self.n_fluid
        pfs = [torch.zeros((nf, ns, kk.size), dtype=torch.float64, device=dev) for kk in allk] if nf else None
        cos = self._cosmo_array(plans, pts, kdev, dt, pfs=pfs)
        cos_dev = self._struct_array(torch, dev, cos, dt)
        # work items: all (cosmology, k), largest k first
        flat = [(-allk[ci][ik], ci, ik) for ci in range(len(plans)) for ik in range(allk[ci].size)]
        flat.sort()
        n_items = len(flat)
        items = [abi.AbcmbPertItem(k=allk[ci][ik], lna_start=plans[ci].lna_start[ik], rtol=plans[ci].rtol[ik],
                                   atol=plans[ci].atol[ik], cosmo=ci, ik=int(ik)) for _, ci, ik in flat]
        items_dev = self._struct_array(torch, dev, items, dt)
        status = torch.zeros(n_items, dtype=torch.int32, device=dev)
        steps = torch.zeros(n_items, dtype=torch.int32, device=dev)
        iters = torch.zeros(n_items, dtype=torch.int32, device=dev)
        state = None
        if want_state:
            state = torch.zeros((n_items, ns, self.layout.n), dtype=torch.float64, device=dev)
        dpts = None
        if tangents is not None:
            nd = abi.N_TANGENT_DIRS
            dpts = [torch.empty((nd, abi.N_PT_FIELDS, ns, kk.size), dtype=torch.float64, device=dev)
                    for kk in allk]
            dpfs = [torch.zeros((nd, nf, ns, kk.size), dtype=torch.float64, device=dev) for kk in allk] if nf \
                else [None] * len(allk)
            keep.append(dpfs)
            tans = [abi.AbcmbTangent(par=dt(sd.par).data_ptr(), tau_tab=dt(sd.tau).data_ptr(),
                                     xe_tab=dt(sd.xe).data_ptr(), tm_tab=dt(sd.tm).data_ptr(), pt=dp.data_ptr(),
                                     pt_fluid=dq.data_ptr() if dq is not None else None)
                    for sd, dp, dq in zip(tangents, dpts, dpfs)]
            tan_dev = self._struct_array(torch, dev, tans, dt)
        nd = abi.N_TANGENT_DIRS if tangents is not None else 0
        ws_bytes = lib.abcmb_pert_workspace_size(len(plans), ns, nd)
        ws = torch.empty(ws_bytes // 8 + 1, dtype=torch.float64, device=dev)
        stage.flush()
        ev0 = ev1 = None
        if self.timing is not None:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
        st = self.settings
        if tangents is None:
            abi.check(lib.abcmb_pert_fwd(cos_dev.data_ptr(), len(plans), items_dev.data_ptr(), n_items, C.byref(st), ns,
                                         status.data_ptr(), steps.data_ptr(), iters.data_ptr(),
                                         state.data_ptr() if state is not None else None, ws.data_ptr(),
                                         ws.numel() * 8, stream),
                      "abcmb_pert_fwd")
        else:
            abi.check(lib.abcmb_pert_tangent(cos_dev.data_ptr(), tan_dev.data_ptr(), len(plans), items_dev.data_ptr(),
                                             n_items, C.byref(st), ns, status.data_ptr(), steps.data_ptr(),
                                             iters.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream),
                      "abcmb_pert_tangent")
        if ev1 is not None:
            ev1.record()
        done = torch.cuda.Event()
        done.record()
        return dict(pts=pts, pfs=pfs, dpts=dpts, kdev=kdev, allk=allk, state=state, flat=flat, keep=keep,
                    status=status,
                    steps=steps, iters=iters, ws=ws, items_dev=items_dev, cos_dev=cos_dev, ev=(ev0, ev1), done=done,
                    tangent=tangents is not None, tan_dev=tan_dev if tangents is not None else None)

    def _finish_ode(self, h):
        """Wait for a launch of ``_launch_ode``, check its status and return the
        tables with the per-k step counts."""
        run = self._ode_result(h)
        run["steps"] = self._check_ode(h)
        return run

    @staticmethod
    def _ode_result(h):
        """The tables of a ``_launch_ode`` handle without waiting for them: work
        enqueued on the launch's stream after this sees them complete."""
        return dict(pts=h["pts"], pfs=h["pfs"], dpts=h["dpts"], kdev=h["kdev"], allk=h["allk"], steps=None,
                    state=h["state"], flat=h["flat"], keep=h["keep"])

    def _check_ode(self, h):
        """Wait for a ``_launch_ode`` handle and raise on any failed (cosmology, k)
        item; returns the per-k step counts of each cosmology."""
        torch = _torch()
        h["done"].synchronize()
        flat, allk, n_items = h["flat"], h["allk"], len(h["flat"])
        st_h = h["status"].cpu().numpy()
        steps_h = h["steps"].cpu().numpy()
        ev0, ev1 = h["ev"]
        if self.timing is not None and ev0 is not None:
            self.timing.append(dict(ms=ev0.elapsed_time(ev1), steps=int(steps_h.sum()),
                                    iters=int(h["iters"].sum().item()), items=n_items,
                                    kernel="tangent" if h["tangent"] else "primal"))
        bad = np.nonzero(st_h != abi_k_ok())[0]
        if bad.size:
            where = [(flat[b][1], float(allk[flat[b][1]][flat[b][2]])) for b in bad[:8]]
            raise RuntimeError(f"perturbation ODE failed for {bad.size} (cosmology, k) items; status codes "
                               f"{np.unique(st_h[bad]).tolist()} (1 = max_steps_PE reached, 2 = non-finite, "
                               f"3 = step-size underflow, 4 = bad input); first (cosmology, k): {where}")
        steps_ck = [np.zeros(a.size, dtype=np.int32) for a in allk]
        for m, (_, ci, ik) in enumerate(flat):
            steps_ck[ci][ik] = steps_h[m]
        return steps_ck

    def _downstream(self, plans, pts, kdev):
        """Sources, LOS C_l, l-spline, lensing and P(k) for ``plans`` with tables ``pts``.

        Runs in chunks of ``self.chunk`` cosmologies.  Returns [B][3][n_ell]
        sampled C_l, [B][3][n_l] C_l on every l, and when lensing is on
        [B][n_l] C_l^phiphi and [B][4][n_l] lensed TT, TE, EE, BB; [B][nq] P(k).
        """
        torch = _torch()
        dev = torch.device("cuda")
        stage = _Stage(torch, dev)
        nkmax = max(int(k.numel()) for k in kdev)
        inp = self._downstream_inputs(plans, pts, kdev, stage.add)
        buf = self._downstream_buffers(len(plans), nkmax)
        stage.flush()
        self._downstream_launch(inp, buf, len(plans), nkmax)
        return {key: buf[key] for key in ("cl_raw", "cl_full", "clpp", "lensed", "pk")}

    def _downstream_inputs(self, plans, pts, kdev, dt):
        """Device inputs of the downstream: the AbcmbCosmo array, the LOS background
        [B][6][n_save-1] and the Limber axes [B][4][n_lens]."""
        torch = _torch()
        dev = torch.device("cuda")
        ncos = len(plans)
        inp = dict(cos=self._struct_array(torch, dev, self._cosmo_array(plans, pts, kdev, dt, tables=False), dt),
                   bgv=dt(np.stack([pl.bgv for pl in plans])), lens=None)
        if self.specs["lensing"]:
            n_lens = plans[0].lens_lna.size
            lens_h = np.zeros((ncos, 4, n_lens))
            for i, pl in enumerate(plans):
                lens_h[i, 0], lens_h[i, 1], lens_h[i, 2] = pl.lens_lna, pl.lens_aH, pl.lens_tau
                lens_h[i, 3, 0] = pl.chi_rec
            inp["lens"] = dt(lens_h)
        return inp

    def _downstream_buffers(self, ncos, nkmax):
        """Outputs and workspaces of the downstream of ``ncos`` cosmologies of at most ``nkmax`` k-modes."""
        torch = _torch()
        lib = self._lib()
        dev = torch.device("cuda")
        f64 = dict(dtype=torch.float64, device=dev)
        ns, nkt, n_ell = self.n_save, self.k_transfer.size, self.ell_idx.size
        nl, nq = self.lensing_ells.size, self.k_pk.size
        chunk = max(1, min(ncos, self.chunk))
        buf = dict(cl_raw=torch.empty((ncos, 3, n_ell), **f64), cl_full=torch.empty((ncos, 3, nl), **f64),
                   pk=torch.empty((ncos, nq), **f64), clpp=None, lensed=None,
                   ws_src=torch.empty(lib.abcmb_sources_workspace_size(chunk, nkmax, ns) // 8 + 1, **f64),
                   ws_los=torch.empty(lib.abcmb_los_workspace_size(chunk, n_ell, nkt) // 8 + 1, **f64),
                   src=torch.empty((chunk, 4, ns - 1, nkt), **f64), ws_lens=None)
        if self.specs["lensing"]:
            buf["clpp"] = torch.empty((ncos, nl), **f64)
            buf["lensed"] = torch.empty((ncos, 4, nl), **f64)
            buf["ws_lens"] = torch.empty(self._lens_ws_bytes(lib, chunk, nl) // 8 + 1, **f64)
        return buf

    def _downstream_launch(self, inp, buf, ncos, nkmax):
        """The downstream kernels on the current stream (no allocation, no copy:
        this sequence is what a HIP graph captures)."""
        torch = _torch()
        lib = self._lib()
        D = self._device_constants()
        stream = torch.cuda.current_stream().cuda_stream
        ns, nkt, n_ell = self.n_save, self.k_transfer.size, self.ell_idx.size
        n_tau, nl, nq = ns - 1, self.lensing_ells.size, self.k_pk.size
        lensing = bool(self.specs["lensing"])
        sc = self.scales.ctypes.data_as(C.c_void_p)
        csz = C.sizeof(abi.AbcmbCosmo)
        chunk = max(1, min(ncos, self.chunk))
        cos_dev, bgv_all, lens_all = inp["cos"], inp["bgv"], inp["lens"]
        n_lens = lens_all.shape[2] if lensing else 0
        cl_raw, cl_full, clpp, lensed, pk = buf["cl_raw"], buf["cl_full"], buf["clpp"], buf["lensed"], buf["pk"]
        ws_src, ws_los, src, ws_lens = buf["ws_src"], buf["ws_los"], buf["src"], buf["ws_lens"]
        for c0 in range(0, ncos, chunk):
            nb = min(chunk, ncos - c0)
            cp = cos_dev.data_ptr() + c0 * csz
            abi.check(lib.abcmb_sources(cp, nb, nkmax, ns, bgv_all[c0].data_ptr(), D["kt"].data_ptr(), nkt, sc,
                                        src.data_ptr(), ws_src.data_ptr(), ws_src.numel() * 8, stream),
                      "abcmb_sources")
            abi.check(lib.abcmb_los_cl(cp, nb, src.data_ptr(), n_tau, bgv_all[c0].data_ptr(), D["kt"].data_ptr(), nkt,
                                       D["bes"].data_ptr(), D["rng"].data_ptr(), D["bes_l"].data_ptr(),
                                       self.bes_l.size, self.bes_tab.shape[2], D["ell_idx"].data_ptr(), n_ell,
                                       self.los_precision, cl_raw[c0].data_ptr(), None, ws_los.data_ptr(),
                                       ws_los.numel() * 8, stream), "abcmb_los_cl")
            abi.check(lib.abcmb_apply_operator(D["ell_op"].data_ptr(), nl, n_ell, cl_raw[c0].data_ptr(), 3 * nb,
                                               cl_full[c0].data_ptr(), stream), "abcmb_apply_operator")
            if lensing:
                abi.check(lib.abcmb_lensing_clpp(cp, nb, lens_all[c0].data_ptr(), n_lens,
                                                 D["lens_ells"].data_ptr(), nl, clpp[c0].data_ptr(), stream),
                          "abcmb_lensing_clpp")
                abi.check(self._lens_call(lib, D, nb, cl_full[c0].data_ptr(), clpp[c0].data_ptr(), nl,
                                          lensed[c0].data_ptr(), ws_lens, stream), "lensing")
            abi.check(lib.abcmb_pk_lin(cp, nb, D["k_pk"].data_ptr(), nq, pk[c0].data_ptr(), stream), "abcmb_pk_lin")

    def downstream_graph(self, n_cosmo: int, nk_max: int):
        """A ``DownstreamGraph`` for batches of exactly ``n_cosmo`` cosmologies of at
        most ``nk_max`` k-modes (built once per shape and kept)."""
        key = (n_cosmo, nk_max)
        cache = self.__dict__.setdefault("_graphs", {})
        if key not in cache:
            cache[key] = DownstreamGraph(self, n_cosmo, nk_max)
        return cache[key]

    def _lens_ws_bytes(self, lib, n_cosmo, nl):
        if self.lens_mode == "mfma":
            return lib.abcmb_lens_mfma_workspace_size(n_cosmo, self.mu.size, nl, self.lens_precision)
        return lib.abcmb_lens_workspace_size(n_cosmo, self.mu.size, nl)

    def _lens_call(self, lib, D, nb, cl_ptr, clpp_ptr, nl, out_ptr, ws, stream):
        """One lensing call on ``nb`` cosmologies through the selected entry point."""
        if self.lens_mode == "mfma":
            return lib.abcmb_lens_mfma(nb, cl_ptr, clpp_ptr, nl, D["mu"].data_ptr(), D["w"].data_ptr(), self.mu.size,
                                       self._dmat().data_ptr(), self.lens_precision, out_ptr, ws.data_ptr(),
                                       ws.numel() * 8, stream)
        return lib.abcmb_lens(nb, cl_ptr, clpp_ptr, nl, D["mu"].data_ptr(), D["w"].data_ptr(), self.mu.size, out_ptr,
                              ws.data_ptr(), ws.numel() * 8, stream)

    def _timed(self, name, work, fn):
        """Run ``fn`` (one C-ABI call on the current stream), bracketed by HIP
        events on that stream when ``call_timing`` is set."""
        if self.call_timing is None:
            return fn()
        torch = _torch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = fn()
        e1.record()
        self.call_timing.setdefault(name, []).append((e0, e1, work))
        return rc

    def _staging(self):
        torch = _torch()
        dev = torch.device("cuda")
        return torch, dev, _Stage(torch, dev)

    def _run_sources(self, plans, pts, kdev):
        """LOS source functions [B][4][n_save-1][n_kt] of ``plans`` with tables ``pts``."""
        torch, dev, stage = self._staging()
        dt = stage.add
        lib = self._lib()
        D = self._device_constants()
        stream = torch.cuda.current_stream().cuda_stream
        ns, ncos, nkt = self.n_save, len(plans), self.k_transfer.size
        nkmax = max(int(k.numel()) for k in kdev)
        cos_dev = self._struct_array(torch, dev, self._cosmo_array(plans, pts, kdev, dt, tables=False), dt)
        bgv_all = dt(np.stack([pl.bgv for pl in plans]))
        src = torch.empty((ncos, 4, ns - 1, nkt), dtype=torch.float64, device=dev)
        chunk = max(1, min(ncos, self.chunk))
        ws = torch.empty(lib.abcmb_sources_workspace_size(chunk, nkmax, ns) // 8 + 1, dtype=torch.float64, device=dev)
        csz = C.sizeof(abi.AbcmbCosmo)
        sc = self.scales.ctypes.data_as(C.c_void_p)
        stage.flush()
        for c0 in range(0, ncos, chunk):
            nb = min(chunk, ncos - c0)
            abi.check(lib.abcmb_sources(cos_dev.data_ptr() + c0 * csz, nb, nkmax, ns, bgv_all[c0].data_ptr(),
                                        D["kt"].data_ptr(), nkt, sc, src[c0].data_ptr(), ws.data_ptr(),
                                        ws.numel() * 8, stream), "abcmb_sources")
        return src

    def downstream_tangent(self, plans, pts, kdev, dpts, plus, minus, steps):
        """Unlensed C_l of ``plans`` and their derivatives, on every l.

        Along the tangent directions d the tables move along dpts[i][d] and the
        host inputs are differenced between plus[i][d] and minus[i][d] (plans at
        theta +- steps[i][d], grids frozen): the sources' derivatives by central
        differences of the (linear-in-the-table) source functions, the
        distances' from the LOS background; the LOS itself is differentiated
        in one tangent-linear pass (``abcmb_los_tangent``).  Returns
        [B][1 + N_TANGENT_DIRS + 2][3][n_l]: value, the directions, d/d ln A_s,
        d/d n_s.
        """
        torch, dev, stage = self._staging()
        dt = stage.add
        lib = self._lib()
        D = self._device_constants()
        stream = torch.cuda.current_stream().cuda_stream
        nd = abi.N_TANGENT_DIRS
        ncos, nkt, n_ell, nl = len(plans), self.k_transfer.size, self.ell_idx.size, self.lensing_ells.size
        n_tau = self.n_save - 1
        nv = 1 + nd + 2
        out = torch.empty((ncos, nv, 3, nl), dtype=torch.float64, device=dev)
        group = max(1, self.chunk // (1 + 2 * nd))
        per = 1 + 2 * nd
        for c0 in range(0, ncos, group):
            cs = range(c0, min(ncos, c0 + group))
            nb = len(cs)
            hs = np.array([[[steps[c][d], -steps[c][d]] for d in range(nd)] for c in cs])   # [nb][nd][2]
            hs_d = dt(hs)
            h2_d = dt(2.0 * hs[:, :, 0])
            dchi = np.zeros((nb, nd, n_tau))
            for i, c in enumerate(cs):
                for d in range(nd):
                    pp, pm = plus[c][d], minus[c][d]
                    dchi[i, d] = ((pp.bg.tau0 - pp.bgv[0]) - (pm.bg.tau0 - pm.bgv[0])) / (2.0 * steps[c][d])
            cos_b = [plans[c] for c in cs]
            cos_dev = self._struct_array(torch, dev, self._cosmo_array(cos_b, [pts[c] for c in cs],
                                                                       [kdev[c] for c in cs], dt, tables=False), dt)
            bgv_b = dt(np.stack([pl.bgv for pl in cos_b]))
            dchi_d = dt(dchi)
            stage.flush()
            vplans, vpts, vk = [], [], []
            for i, c in enumerate(cs):
                vplans.append(plans[c]); vpts.append(pts[c]); vk.append(kdev[c])
                moved = _moved_tables(torch, pts[c], dpts[c], hs_d[i])
                for d in range(nd):
                    vplans += [plus[c][d], minus[c][d]]
                    vpts += [moved[2 * d], moved[2 * d + 1]]
                    vk += [kdev[c], kdev[c]]
            src = self._run_sources(vplans, vpts, vk)
            del vpts, moved
            sv = src.view(nb, per, 4, n_tau, nkt)
            src_b = sv[:, 0].contiguous()
            # central differences of the sources: (S(+h) - S(-h)) / 2h
            dsrc = torch.sub(sv[:, 1::2], sv[:, 2::2]).div_(h2_d.view(nb, nd, 1, 1, 1))
            del src, sv
            ws = torch.empty(lib.abcmb_los_tangent_workspace_size(nb, nd, n_ell, nkt) // 8 + 1, dtype=torch.float64,
                             device=dev)
            raw = torch.empty((nb, nv, 3, n_ell), dtype=torch.float64, device=dev)
            abi.check(self._timed("abcmb_los_tangent", dict(n_cosmo=nb, n_ell=n_ell, nkt=nkt, n_tau=n_tau, n_dir=nd),
                                  lambda: lib.abcmb_los_tangent(
                                      cos_dev.data_ptr(), nb, src_b.data_ptr(), dsrc.data_ptr(), dchi_d.data_ptr(),
                                      nd, n_tau, bgv_b.data_ptr(), D["kt"].data_ptr(), nkt, D["bes"].data_ptr(),
                                      D["rng"].data_ptr(), D["bes_l"].data_ptr(), self.bes_l.size,
                                      self.bes_tab.shape[2], D["ell_idx"].data_ptr(), n_ell, self.los_precision,
                                      raw.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream)), "abcmb_los_tangent")
            abi.check(lib.abcmb_apply_operator(D["ell_op"].data_ptr(), nl, n_ell, raw.data_ptr(), nb * nv * 3,
                                               out[c0].data_ptr(), stream), "abcmb_apply_operator")
        return out

    def downstream_adjoint(self, plans, pts, kdev, ds, obs_bar):
        """Reverse mode of ``_downstream`` for ``plans`` with tables ``pts``.

        ``ds``: the forward result of ``_downstream`` on the same inputs;
        ``obs_bar``: [B][3][n_l] cotangents of ``observed_cl(ds)``.  The lensing
        (``abcmb_lens_adjoint``), the l-spline (its operator transposed), the
        sources, LOS and k-quadrature (``abcmb_los_adjoint``) and the table part
        of C_l^phiphi (``abcmb_lensing_clpp_adjoint``) run backwards on the GPU.
        Returns dict(pt_bar=[21][n_save][nk] per plan (dL/d table), bgv_bar
        [B][6][n_save-1] (dL/d LOS background vectors), scal_bar [B][3] (dL/d
        A_s, n_s, tau0 through the unlensed C_l), clpp_bar [B][n_l] (dL/d
        C_l^phiphi; None without lensing)).
        """
        torch, dev, stage = self._staging()
        dt = stage.add
        lib = self._lib()
        D = self._device_constants()
        stream = torch.cuda.current_stream().cuda_stream
        f64 = dict(dtype=torch.float64, device=dev)
        ncos, nl, n_l = len(plans), self.lensing_ells.size, self.ells.size
        ns, nkt, n_ell = self.n_save, self.k_transfer.size, self.ell_idx.size
        n_tau = ns - 1
        nkmax = max(int(k.numel()) for k in kdev)
        lensing = bool(self.specs["lensing"])
        inp = self._downstream_inputs(plans, pts, kdev, dt)
        pt_bar = [torch.zeros_like(pt) for pt in pts]
        ptrs = dt(np.array([p.data_ptr() for p in pt_bar], dtype=np.int64), torch.int64)
        if "ell_op_t" not in D:
            D["ell_op_t"] = torch.as_tensor(np.ascontiguousarray(self.ell_op.T), **f64)
        stage.flush()
        obs_bar = obs_bar.to(torch.float64).contiguous()
        cl_full_bar = torch.zeros((ncos, 3, nl), **f64)
        clpp_bar = None
        csz = C.sizeof(abi.AbcmbCosmo)
        chunk = max(1, min(ncos, self.chunk))
        cos_dev = inp["cos"]
        if lensing:
            lensed_bar = torch.zeros((ncos, 4, nl), **f64)
            lensed_bar[:, :3, :n_l] = obs_bar
            clpp_bar = torch.empty((ncos, nl), **f64)
            ws = torch.empty(lib.abcmb_lens_adjoint_workspace_size(chunk, self.mu.size, nl) // 8 + 1, **f64)
            cl_full, clpp = ds["cl_full"].contiguous(), ds["clpp"].contiguous()
            n_lens = inp["lens"].shape[2]
            for c0 in range(0, ncos, chunk):
                nb = min(chunk, ncos - c0)
                abi.check(lib.abcmb_lens_adjoint(nb, cl_full[c0].data_ptr(), clpp[c0].data_ptr(), nl,
                                                 D["mu"].data_ptr(), D["w"].data_ptr(), self.mu.size,
                                                 lensed_bar[c0].data_ptr(), cl_full_bar[c0].data_ptr(),
                                                 clpp_bar[c0].data_ptr(), ws.data_ptr(), ws.numel() * 8, stream),
                          "abcmb_lens_adjoint")
            del ws
            abi.check(lib.abcmb_lensing_clpp_adjoint(cos_dev.data_ptr(), ncos, inp["lens"].data_ptr(), n_lens,
                                                     D["lens_ells"].data_ptr(), nl, clpp_bar.data_ptr(),
                                                     ptrs.data_ptr(), stream), "abcmb_lensing_clpp_adjoint")
        else:
            cl_full_bar[:, :, :n_l] = obs_bar
        cl_raw_bar = torch.empty((ncos, 3, n_ell), **f64)
        abi.check(lib.abcmb_apply_operator(D["ell_op_t"].data_ptr(), n_ell, nl, cl_full_bar.data_ptr(), 3 * ncos,
                                           cl_raw_bar.data_ptr(), stream), "abcmb_apply_operator")
        bgv_bar = torch.empty((ncos, 6, n_tau), **f64)
        scal_bar = torch.empty((ncos, 3), **f64)
        # the LOS adjoint holds ~10 n_save (n_kt + nk) doubles per cosmology: groups of a few
        per = lib.abcmb_los_adjoint_workspace_size(1, nkmax, ns, n_ell, nkt)
        group = int(max(1, min(chunk, (4 << 30) // max(per, 1))))
        ws = torch.empty(lib.abcmb_los_adjoint_workspace_size(group, nkmax, ns, n_ell, nkt) // 8 + 1, **f64)
        sc = self.scales.ctypes.data_as(C.c_void_p)
        for c0 in range(0, ncos, group):
            nb = min(group, ncos - c0)
            abi.check(lib.abcmb_los_adjoint(cos_dev.data_ptr() + c0 * csz, nb, nkmax, ns, inp["bgv"][c0].data_ptr(),
                                            D["kt"].data_ptr(), nkt, sc, D["bes"].data_ptr(), D["rng"].data_ptr(),
                                            D["bes_l"].data_ptr(), self.bes_l.size, self.bes_tab.shape[2],
                                            D["ell_idx"].data_ptr(), n_ell, cl_raw_bar[c0].data_ptr(),
                                            ptrs[c0:].data_ptr(), bgv_bar[c0].data_ptr(), scal_bar[c0].data_ptr(),
                                            ws.data_ptr(), ws.numel() * 8, stream), "abcmb_los_adjoint")
        del ws
        return dict(pt_bar=pt_bar, bgv_bar=bgv_bar, scal_bar=scal_bar, clpp_bar=clpp_bar, keep=(stage, ptrs, inp))

    def clpp_of(self, plans, pts, kdev):
        """C_l^phiphi [B][n_l] of ``plans`` with the delta_m fields of tables ``pts``
        (any plan may share a table: the host inputs move, the table does not)."""
        torch, dev, stage = self._staging()
        dt = stage.add
        lib = self._lib()
        D = self._device_constants()
        stream = torch.cuda.current_stream().cuda_stream
        ncos, nl = len(plans), self.lensing_ells.size
        cos_dev = self._struct_array(torch, dev, self._cosmo_array(plans, pts, kdev, dt, tables=False), dt)
        n_lens = plans[0].lens_lna.size
        lens_h = np.zeros((ncos, 4, n_lens))
        for i, pl in enumerate(plans):
            lens_h[i, 0], lens_h[i, 1], lens_h[i, 2] = pl.lens_lna, pl.lens_aH, pl.lens_tau
            lens_h[i, 3, 0] = pl.chi_rec
        lens_all = dt(lens_h)
        clpp = torch.empty((ncos, nl), dtype=torch.float64, device=dev)
        stage.flush()
        abi.check(lib.abcmb_lensing_clpp(cos_dev.data_ptr(), ncos, lens_all.data_ptr(), n_lens,
                                         D["lens_ells"].data_ptr(), nl, clpp.data_ptr(), stream),
                  "abcmb_lensing_clpp")
        torch.cuda.current_stream().synchronize()
        return clpp

    def lensed_from_cl(self, plans, pts, kdev, cl_full):
        """Lensed [B][4][n_l] from unlensed cl_full [B][3][n_l], with C_l^phiphi of ``plans`` / ``pts``."""
        torch, dev, stage = self._staging()
        dt = stage.add
        lib = self._lib()
        D = self._device_constants()
        stream = torch.cuda.current_stream().cuda_stream
        ncos, nl = len(plans), self.lensing_ells.size
        cos_dev = self._struct_array(torch, dev, self._cosmo_array(plans, pts, kdev, dt, tables=False), dt)
        n_lens = plans[0].lens_lna.size
        lens_h = np.zeros((ncos, 4, n_lens))
        for i, pl in enumerate(plans):
            lens_h[i, 0], lens_h[i, 1], lens_h[i, 2] = pl.lens_lna, pl.lens_aH, pl.lens_tau
            lens_h[i, 3, 0] = pl.chi_rec
        lens_all = dt(lens_h)
        cl_full = cl_full.contiguous()
        clpp = torch.empty((ncos, nl), dtype=torch.float64, device=dev)
        lensed = torch.empty((ncos, 4, nl), dtype=torch.float64, device=dev)
        chunk = max(1, min(ncos, self.chunk))
        ws = torch.empty(self._lens_ws_bytes(lib, chunk, nl) // 8 + 1, dtype=torch.float64, device=dev)
        csz = C.sizeof(abi.AbcmbCosmo)
        stage.flush()
        for c0 in range(0, ncos, chunk):
            nb = min(chunk, ncos - c0)
            abi.check(lib.abcmb_lensing_clpp(cos_dev.data_ptr() + c0 * csz, nb, lens_all[c0].data_ptr(), n_lens,
                                             D["lens_ells"].data_ptr(), nl, clpp[c0].data_ptr(), stream),
                      "abcmb_lensing_clpp")
            abi.check(self._timed("lensing", dict(n_cosmo=nb, n_mu=self.mu.size, n_l=nl, mode=self.lens_mode),
                                  lambda: self._lens_call(lib, D, nb, cl_full[c0].data_ptr(), clpp[c0].data_ptr(), nl,
                                                          lensed[c0].data_ptr(), ws, stream)), "lensing")
        return lensed

    def _pk(self, k, z, PT, params, kind):
        torch = _torch()
        lib = self._lib()
        dev = torch.device("cuda")
        _, _, stage = self._staging()
        kq = stage.add(np.atleast_1d(np.asarray(k, dtype=np.float64)))
        lna = stage.add(np.asarray(PT.lna, dtype=np.float64))
        kd = getattr(PT, "k_dev", None)
        if kd is None:
            kd = stage.add(np.asarray(PT.k, dtype=np.float64))
        table = PT.table if hasattr(PT.table, "data_ptr") else stage.add(np.asarray(PT.table))
        c = abi.AbcmbCosmo()
        c.A_s, c.n_s, c.k_pivot = params["A_s"], params["n_s"], self.specs["k_pivot"]
        c.omega_b, c.omega_cdm, c.omega_m = params["omega_b"], params["omega_cdm"], params["omega_m"]
        c.n_save, c.nk = int(lna.numel()), int(kd.numel())
        c.lna_save, c.k, c.pt = lna.data_ptr(), kd.data_ptr(), table.data_ptr()
        cos = self._struct_array(torch, dev, [c], stage.add)
        out = torch.empty(kq.numel(), dtype=torch.float64, device=dev)
        stage.flush()
        abi.check(lib.abcmb_pk(cos.data_ptr(), 1, kq.data_ptr(), int(kq.numel()), float(z), kind, out.data_ptr(),
                               torch.cuda.current_stream().cuda_stream), "abcmb_pk")
        return out

    def Pk_lin(self, k, z, PT, params):
        """Linear matter power spectrum P(k, z) in Mpc^3 on the GPU (SpectrumSolver.Pk_lin,
        abcmb/spectrum.py:268-302)."""
        return self._pk(k, z, PT, params, abi.PK_M)

    def Pk_cb(self, k, z, PT, params):
        """Linear CDM + baryon power spectrum P_cb(k, z) in Mpc^3 on the GPU
        (SpectrumSolver.Pk_cb, abcmb/spectrum.py:304-347)."""
        return self._pk(k, z, PT, params, abi.PK_CB)

    @property
    def SS(self):
        """The reference's spectrum-solver handle (``model.SS.Pk_lin(...)``): the
        P(k) entry points live on the model here."""
        return self

    def observed_cl(self, ds):
        """[B][3][n_l] TT, TE, EE on l = 2..l_max (lensed when lensing is on) from a downstream result."""
        n_l = self.ells.size
        if self.specs["lensing"]:
            return ds["lensed"][:, :3, :n_l]
        return ds["cl_full"][:, :, :n_l]

    def _outputs(self, plans, run, ds):
        n_l = self.ells.size
        lensing = bool(self.specs["lensing"])
        outs = []
        for i, pl in enumerate(plans):
            if lensing:
                L = ds["lensed"]
                TT, TE, EE, BB = L[i, 0, :n_l], L[i, 1, :n_l], L[i, 2, :n_l], L[i, 3, :n_l]
            else:
                F = ds["cl_full"]
                TT, TE, EE, BB = F[i, 0, :n_l], F[i, 1, :n_l], F[i, 2, :n_l], None
            PT = PerturbationTable(run["allk"][i], pl.lna_save, run["pts"][i], self.has_ncdm, self.species_off,
                                   self.fluids, run["pfs"][i] if run.get("pfs") else None, (pl.bg, pl.params))
            PT.k_dev = run["kdev"][i]
            out = Output(TT, TE, EE, ds["pk"][i], self.ells, self.k_pk, pl.bg, PT, pl.params, ClBB=BB,
                         Clpp=ds["clpp"][i] if lensing else None,
                         steps=run["steps"][i] if run["steps"] is not None else None)
            out.cl_raw = ds["cl_raw"][i]
            out.cl_unlensed = ds["cl_full"][i]
            outs.append(out)
        return outs

    def execute(self, plans, want_state: bool = False):
        """Outputs of prepared cosmologies: one ODE launch, then the batched downstream kernels."""
        run = self._run_ode(plans, want_state=want_state)
        ds = self._downstream(plans, run["pts"], run["kdev"])
        outs = self._outputs(plans, run, ds)
        if want_state:
            return outs, run["state"], run["flat"]
        return outs

    def cl_jacobian(self, params: dict, names, rel_step: float = 1e-4) -> dict:
        """Unlensed C_l of one cosmology and their derivatives along the
        parameters ``names`` -- input parameters (of ``params`` or of
        ``fix_params``) that reach the perturbation equations, the user fluids'
        own parameters included (what ``jax.jacfwd`` of the reference's
        ``Model.__call__`` gives per parameter), at most N_TANGENT_DIRS of them
        -- plus d/d ln A_s and d/d n_s.  The host inputs' derivatives come from the producer's forward
        mode (``tangent_plans``, the derived parameters by central differences
        at a relative step ``rel_step``), the tables' from one tangent-linear ODE
        launch, the spectra's from the tangent-linear LOS.  Returns
        ``{"l": l, "cl": [3][n_l] TT, TE, EE, name: [3][n_l] d C_l / d name,
        "ln10As" and "n_s": [3][n_l]}`` (numpy, l = 2..l_max); with lensing also
        ``"lensed"`` ([4][n_l] TT, TE, EE, BB) and ``"lensed_" + name``, the
        lensing differenced at C_l +- h dC_l as ``tangent_gradients`` does."""
        names = list(names)
        nd = abi.N_TANGENT_DIRS
        if not names or len(names) > nd or len(set(names)) != len(names):
            raise ValueError(f"cl_jacobian: between 1 and {nd} distinct parameter names")
        if any(n in ("A_s", "ln10As", "n_s") for n in names):
            raise ValueError("cl_jacobian: A_s and n_s do not reach the perturbations; "
                             "their derivatives are always returned")
        full = self.derived(params)
        for n in names:
            if n not in full:
                raise KeyError(f"cl_jacobian: {n!r} is not a parameter of this model")
        base = self.prepare(full)
        steps, pm_params = [], []
        for n in names:
            h = rel_step * abs(float(full[n])) or rel_step
            steps.append(h)
            src = dict(self.fixed_params, **params)
            for sgn in (1.0, -1.0):
                q = dict(src)
                q[n] = float(full[n]) + sgn * h
                pm_params.append(self.derived(q))
        pm = self.tangent_plans(base, pm_params, steps)
        plus, minus = pm[0::2], pm[1::2]
        seeds = tangent_seeds(plus, minus, steps)
        h = self.launch_tangent([base], [seeds])
        run = self._ode_result(h)
        # directions beyond ``names`` carry zero seeds: their plans are the base
        # plan itself, at a unit step
        pad = nd - len(names)
        cl_t = self.downstream_tangent([base], run["pts"], run["kdev"], run["dpts"], [plus + [base] * pad],
                                       [minus + [base] * pad], [steps + [1.0] * pad])
        n_l = self.ells.size
        lens = None
        if self.specs["lensing"]:
            # lensed spectra at theta +- h along each name: the lensing of
            # C_l +- h dC_l with C_l^phiphi of the tables moved along their
            # tangents and the plans at theta +- h; central differences
            torch = _torch()
            pt0, kd0 = run["pts"][0], run["kdev"][0]
            hs = torch.tensor([[st, -st] for st in steps], dtype=torch.float64, device=pt0.device)
            moved = _moved_tables(torch, pt0[:1], run["dpts"][0][:len(names), :1], hs)   # delta_m alone
            vplans, vpts, cls = [base], [pt0[:1]], [cl_t[0, 0]]
            for d in range(len(names)):
                for si, sgn in enumerate((1.0, -1.0)):
                    vplans.append(pm[2 * d + si])
                    vpts.append(moved[2 * d + si])
                    cls.append(cl_t[0, 0] + sgn * steps[d] * cl_t[0, 1 + d])
            lens = self.lensed_from_cl(vplans, vpts, [kd0] * len(vplans), torch.stack(cls))[:, :, :n_l].cpu().numpy()
        self._check_ode(h)
        c = cl_t[0, :, :, :n_l].cpu().numpy()
        out = {"l": np.arange(2, 2 + n_l), "cl": c[0]}
        for i, n in enumerate(names):
            out[n] = c[1 + i]
        out["ln10As"], out["n_s"] = c[1 + nd], c[2 + nd]
        if lens is not None:
            out["lensed"] = lens[0]
            for i, n in enumerate(names):
                out["lensed_" + n] = (lens[1 + 2 * i] - lens[2 + 2 * i]) / (2.0 * steps[i])
        return out

    def execute_tangent(self, plans, seeds):
        """Outputs of ``plans`` plus the derivatives of their perturbation tables.

        ``seeds``: one ``TangentSeeds`` per plan.  Returns (outputs, dpts) with
        dpts[i] of shape [N_TANGENT_DIRS][21][n_save][nk_i] on the GPU.
        """
        return self.finish_tangent(plans, self.launch_tangent(plans, seeds))

    def launch_tangent(self, plans, seeds):
        """Enqueue the tangent-linear ODE launch of ``execute_tangent`` on the
        current stream and return its handle (``_check_ode`` waits for it)."""
        return self._launch_ode(plans, tangents=seeds)

    def finish_tangent(self, plans, handle):
        """Wait for a ``launch_tangent`` handle, run the downstream on the current
        stream and return (outputs, dpts) as ``execute_tangent`` does."""
        run = self._finish_ode(handle)
        ds = self._downstream(plans, run["pts"], run["kdev"])
        return self._outputs(plans, run, ds), run["dpts"]

    def ode_adjoint(self, plans, pt_bars):
        """Reverse-mode cotangents of the host inputs of ``plans``' perturbation tables.

        ``pt_bars[i]``: dL/d(table of plans[i]), a [21][n_save][nk_i] float64 GPU
        tensor.  One ``abcmb_pert_adjoint`` launch for all (cosmology, k) items
        (the same items, order and step sequences as ``execute``).  Returns one
        dict per plan with the GPU tensors ``par`` (dL/d of the 31 AbcmbCosmo
        double fields), ``tau``, ``xe`` and ``tm`` (dL/d of the table entries),
        after checking every item's status.
        """
        self._complete_pending(plans)
        torch = _torch()
        lib = self._lib()
        D = self._device_constants()
        dev = torch.device("cuda")
        stream = torch.cuda.current_stream().cuda_stream
        ns = self.n_save
        stage = _Stage(torch, dev)
        dt = stage.add
        allk = [(pl.k if pl.k is not None else self.k_pt) for pl in plans]
        kdev = [D["k_pt"] if kk is self.k_pt else dt(kk) for kk in allk]
        for pb, kk in zip(pt_bars, allk):
            if tuple(pb.shape) != (abi.N_PT_FIELDS, ns, kk.size) or pb.dtype != torch.float64 or not pb.is_cuda:
                raise ValueError(f"table cotangent of shape {tuple(pb.shape)} / {pb.dtype}; expected "
                                 f"({abi.N_PT_FIELDS}, {ns}, {kk.size}) float64 on the GPU")
        pt_bars = [pb.contiguous() for pb in pt_bars]
        cos = self._cosmo_array(plans, pt_bars, kdev, dt)
        cos_dev = self._struct_array(torch, dev, cos, dt)
        flat = [(-allk[ci][ik], ci, ik) for ci in range(len(plans)) for ik in range(allk[ci].size)]
        flat.sort()
        n_items = len(flat)
        items = [abi.AbcmbPertItem(k=allk[ci][ik], lna_start=plans[ci].lna_start[ik], rtol=plans[ci].rtol[ik],
                                   atol=plans[ci].atol[ik], cosmo=ci, ik=int(ik)) for _, ci, ik in flat]
        items_dev = self._struct_array(torch, dev, items, dt)
        out = []
        for pl in plans:
            out.append(dict(par=torch.zeros(abi.COSMO_DOUBLES, dtype=torch.float64, device=dev),
                            tau=torch.zeros(abi.TAU_TAB_N, dtype=torch.float64, device=dev),
                            xe=torch.zeros(pl.bg.xe_tab.size, dtype=torch.float64, device=dev),
                            tm=torch.zeros(pl.bg.tm_tab.size, dtype=torch.float64, device=dev)))
        cots = [abi.AbcmbCotangent(pt=pb.data_ptr(), par=o["par"].data_ptr(), tau_tab=o["tau"].data_ptr(),
                                   xe_tab=o["xe"].data_ptr(), tm_tab=o["tm"].data_ptr())
                for pb, o in zip(pt_bars, out)]
        cot_dev = self._struct_array(torch, dev, cots, dt)
        status = torch.zeros(n_items, dtype=torch.int32, device=dev)
        steps = torch.zeros(n_items, dtype=torch.int32, device=dev)
        iters = torch.zeros(n_items, dtype=torch.int32, device=dev)
        tab_max = max([abi.TAU_TAB_N] + [max(pl.bg.xe_tab.size, pl.bg.tm_tab.size) for pl in plans])
        st = self.settings
        ws_bytes = lib.abcmb_pert_adjoint_workspace_size(len(plans), ns, tab_max, C.byref(st))
        if ws_bytes == 0:
            raise RuntimeError(f"abcmb_pert_adjoint_workspace_size: {lib.abcmb_last_error().decode()}")
        ws = torch.empty(ws_bytes // 8 + 1, dtype=torch.float64, device=dev)
        stage.flush()
        ev0 = ev1 = None
        if self.timing is not None:
            ev0 = torch.cuda.Event(enable_timing=True)
            ev1 = torch.cuda.Event(enable_timing=True)
            ev0.record()
        abi.check(lib.abcmb_pert_adjoint(cos_dev.data_ptr(), cot_dev.data_ptr(), len(plans), items_dev.data_ptr(),
                                         n_items, C.byref(st), ns, tab_max, status.data_ptr(), steps.data_ptr(),
                                         iters.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream),
                  "abcmb_pert_adjoint")
        if ev1 is not None:
            ev1.record()
        torch.cuda.current_stream().synchronize()
        if ev0 is not None:
            self.timing.append(dict(ms=ev0.elapsed_time(ev1), steps=int(steps.sum().item()),
                                    iters=int(iters.sum().item()), items=n_items, kernel="adjoint"))
        st_h = status.cpu().numpy()
        bad = np.nonzero(st_h != abi_k_ok())[0]
        if bad.size:
            where = [(flat[b][1], float(allk[flat[b][1]][flat[b][2]])) for b in bad[:8]]
            raise RuntimeError(f"perturbation adjoint failed for {bad.size} (cosmology, k) items; status codes "
                               f"{np.unique(st_h[bad]).tolist()} (5 = non-finite cotangent); "
                               f"first (cosmology, k): {where}")
        del stage, ws
        steps_ck = [np.zeros(a.size, dtype=np.int32) for a in allk]
        steps_h = steps.cpu().numpy()
        for m, (_, ci, ik) in enumerate(flat):
            steps_ck[ci][ik] = steps_h[m]
        for o, s_ in zip(out, steps_ck):
            o["steps"] = s_
        return out

    def __call__(self, params: dict = None) -> Output:
        return self.run_batch([params if params is not None else {}])[0]


class DownstreamGraph:
    """The downstream of a batch (sources, LOS, l-spline, C_l^phiphi, lensing,
    P(k)) captured once as a HIP graph and replayed for new inputs.

    Every input lives in a static device buffer (the cosmology structs, the
    LOS and Limber background vectors, one perturbation-table slot and one
    k-grid slot per cosmology); ``run`` copies a batch's inputs into them on
    the current stream and replays the graph after the copies.  The captured
    sequence holds kernels only: no allocation, no copy, no memset.  Results
    are the static output buffers (overwritten by the next ``run``).
    """

    def __init__(self, model, n_cosmo, nk_max):
        torch = _torch()
        dev = torch.device("cuda")
        self.model, self.n_cosmo, self.nk_max = model, n_cosmo, nk_max
        ns = model.n_save
        f64 = dict(dtype=torch.float64, device=dev)
        model._device_constants()
        if model.specs["lensing"] and model.lens_mode == "mfma":
            model._dmat()        # built outside the capture
        self.pt = torch.zeros((n_cosmo, abi.N_PT_FIELDS * ns * nk_max), **f64)
        self.k = torch.ones((n_cosmo, nk_max), **f64)
        self.lna = torch.zeros((n_cosmo, ns), **f64)
        csz = C.sizeof(abi.AbcmbCosmo)
        self.inp = dict(cos=torch.zeros(n_cosmo * csz, dtype=torch.uint8, device=dev),
                        bgv=torch.zeros((n_cosmo, 6, ns - 1), **f64), lens=None)
        if model.specs["lensing"]:
            self.inp["lens"] = torch.zeros((n_cosmo, 4, 500), **f64)
        self.buf = model._downstream_buffers(n_cosmo, nk_max)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            with torch.cuda.graph(self.graph, stream=side):
                model._downstream_launch(self.inp, self.buf, n_cosmo, nk_max)
        torch.cuda.current_stream().wait_stream(side)

    def run(self, plans, pts, kdev):
        """Downstream of ``plans`` (exactly ``n_cosmo`` of them) with tables ``pts``."""
        torch = _torch()
        m = self.model
        if len(plans) != self.n_cosmo:
            raise ValueError(f"this graph holds {self.n_cosmo} cosmologies, got {len(plans)}")
        ns = m.n_save
        stage = _Stage(torch, torch.device("cuda"))
        slots_pt, slots_k = [], []
        for i, (pl, kd) in enumerate(zip(plans, kdev)):
            nk = int(kd.numel())
            if nk > self.nk_max:
                raise ValueError(f"cosmology {i} has {nk} k-modes; this graph holds {self.nk_max}")
            if pl.lens_lna is not None and pl.lens_lna.size != 500:
                raise ValueError("the Limber axis must have 500 points")
            slots_pt.append(self.pt[i, :abi.N_PT_FIELDS * ns * nk])
            slots_k.append(self.k[i, :nk])
        # staged host inputs (their device copies are valid after the flush)
        lna_st = stage.add(np.stack([pl.lna_save for pl in plans]))
        cos = m._cosmo_array(plans, slots_pt, slots_k, stage.add, tables=False)
        for i, c in enumerate(cos):
            c.lna_save = self.lna[i].data_ptr()
        cos_st = m._struct_array(torch, torch.device("cuda"), cos, stage.add)
        vec = m._downstream_inputs(plans, pts, kdev, stage.add)
        stage.flush()
        for sp, sk, pt, kd in zip(slots_pt, slots_k, pts, kdev):
            sp.copy_(pt.reshape(-1))
            sk.copy_(kd)
        self.lna.copy_(lna_st)
        self.inp["cos"].copy_(cos_st)
        self.inp["bgv"].copy_(vec["bgv"])
        if self.inp["lens"] is not None:
            self.inp["lens"].copy_(vec["lens"])
        self.graph.replay()
        return {key: self.buf[key] for key in ("cl_raw", "cl_full", "clpp", "lensed", "pk")}


def _moved_tables(torch, pt, dpt, hs):
    """pt + h dpt for every tangent direction and sign in one launch:
    [2 * n_dir][fields][n_save][nk], direction-major, +h before -h.  ``hs``:
    [n_dir][2] steps on the GPU."""
    nd = dpt.shape[0]
    return torch.addcmul(pt.unsqueeze(0).unsqueeze(0), hs.view(nd, 2, 1, 1, 1), dpt.unsqueeze(1)).view(
        2 * nd, *pt.shape)


def _np_dtypes():
    import torch
    return {torch.float64: np.float64, torch.float32: np.float32, torch.int64: np.int64, torch.int32: np.int32,
            torch.uint8: np.uint8}


class _DtypeMap(dict):
    """torch dtype -> numpy dtype, built on first use (torch imports lazily)."""
    def __missing__(self, key):
        self.update(_np_dtypes())
        return dict.__getitem__(self, key)


_NP_DTYPE = _DtypeMap()


class _Stage:
    """Host-to-device staging that never waits for the GPU.

    ``add`` packs a host array into a pinned host segment and returns the
    device tensor it will occupy (its address is final at once, so it can go
    into the structs the kernels read); ``flush`` enqueues one asynchronous
    copy per segment on the current stream and must come before the first
    kernel that reads the staged arrays.  torch's pinned-memory cache keeps a
    host segment until its copy has run.
    """
    SEG = 8 << 20
    ALIGN = 256

    def __init__(self, torch, dev):
        self.torch, self.dev = torch, dev
        self.segs = []      # [host, device, used bytes, host as numpy, bytes already copied]
        self.dirty = False

    def add(self, a, dtype=None):
        torch = self.torch
        dtype = dtype or torch.float64
        arr = np.ascontiguousarray(a, dtype=_NP_DTYPE[dtype])
        nb = arr.nbytes
        if not self.segs or self.segs[-1][2] + nb > self.segs[-1][0].numel():
            size = max(self.SEG, nb)
            host = torch.empty(size, dtype=torch.uint8, pin_memory=True)
            self.segs.append([host, torch.empty(size, dtype=torch.uint8, device=self.dev), 0, host.numpy(), 0])
        seg = self.segs[-1]
        off = seg[2]
        seg[3][off:off + nb] = arr.reshape(-1).view(np.uint8)     # numpy byte copy: no torch CPU threads
        seg[2] = off + (nb + self.ALIGN - 1) // self.ALIGN * self.ALIGN
        self.dirty = True
        return seg[1][off:off + nb].view(dtype).view(arr.shape)

    def flush(self):
        """Copy what was added since the last flush (each byte goes over once)."""
        if self.dirty:
            for seg in self.segs:
                host, devb, used, done = seg[0], seg[1], seg[2], seg[4]
                n = min(used, host.numel())
                if n > done:
                    devb[done:n].copy_(host[done:n], non_blocking=True)
                    seg[4] = n
            self.dirty = False


def abi_k_ok():
    return 0
