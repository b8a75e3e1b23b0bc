"""Host background: the tables the GPU perturbation/LOS path reads.

``Background`` plays the role of the reference's ``Background`` object
(``abcmb/background.py:420-768``) for this library: it owns the conformal
time table, the x_e / T_m histories (HyRex, ``abcmb/hyrex/*.py``, restated in
``csrc/hyrex.cpp``), tau_0, lna_rec, lna_transfer_start and z_reion, produced
by the native host producer ``libabcmb_host.so``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _set_inputs(c, p: dict, k_pivot: float):
    """The parameter fields of an AbcmbCosmo from derived parameters."""
    c.h, c.omega_b, c.omega_cdm, c.omega_Lambda = p["h"], p["omega_b"], p["omega_cdm"], p["omega_Lambda"]
    c.A_s, c.n_s, c.k_pivot, c.TCMB0, c.YHe = p["A_s"], p["n_s"], k_pivot, p["TCMB0"], p["YHe"]
    c.N_ur, c.T_ur = p["N_nu_massless"], p["T_nu_massless"]
    c.N_ncdm, c.T_ncdm, c.m_ncdm = p["N_nu_massive"], p["T_nu_massive"], p["m_nu_massive"]
    c.R_nu, c.R_b, c.om, c.omega_r, c.omega_m = p["R_nu"], p["R_b"], p["om"], p["omega_r"], p["omega_m"]


def _doubles(c) -> np.ndarray:
    """A writable view of the COSMO_DOUBLES double fields of an AbcmbCosmo."""
    return np.frombuffer((C.c_double * abi.COSMO_DOUBLES).from_buffer(c), dtype=np.float64)


def input_doubles(params: dict, k_pivot: float, fluid_par=None) -> np.ndarray:
    """The COSMO_DOUBLES fields an AbcmbCosmo takes from ``params`` and the
    user fluids' parameter slots (the producer's results left zero)."""
    c = abi.AbcmbCosmo()
    _set_inputs(c, params, k_pivot)
    if fluid_par is not None:
        c.fluid_par[:] = [float(v) for v in fluid_par]
    return _doubles(c).copy()


class ProducerTangents:
    """Derivatives of one background's producer along the tangent directions
    (forward mode, every grid choice held): ``par`` [N_TANGENT_DIRS][COSMO_DOUBLES]
    of the AbcmbCosmo doubles, ``tau``, ``xe``, ``tm`` of the tables."""

    def __init__(self, par, tau, xe, tm):
        self.par, self.tau, self.xe, self.tm = par, tau, xe, tm


class Background:
    """Background of one cosmology (host-resident FP64 tables)."""

    def __init__(self, params: dict, has_ncdm: bool, n_q: int = 3, k_pivot: float = 0.05,
                 input_tau_reion: bool = True, xe_pad: int = 0, rec_switch=None,
                 lna_rec: float = None, lna_transfer_start: float = None, reionization: bool = True,
                 species_off: int = 0, fluid_par=None, lib=None):
        """``rec_switch`` (the recombination stages' grid counts, ``REC_STAGES``
        ints), ``lna_rec``, ``lna_transfer_start``: grid choices held at another
        cosmology's (finite differences with frozen grids); the producer then
        keeps them instead of searching for its own.  ``reionization=False``:
        the x_e table holds the recombination history alone and z_reion is 0
        (``abcmb_reionize`` adds the reionization on the GPU).  ``species_off``: the
        AbcmbCosmo.species_off bits; ``fluid_par``: the user fluids' parameter slots;
        ``lib``: the host producer built for the model's user fluids (default: the
        standard one)."""
        lib = lib if lib is not None else abi.host_lib()
        self.lib = lib
        self.params = params
        n_rec = lib.abcmb_host_recomb_points()
        n_tm = lib.abcmb_host_tm_points()
        self.tau_tab = np.zeros(abi.TAU_TAB_N)
        self.xe_tab = np.zeros(n_rec + xe_pad)
        self.tm_tab = np.zeros(n_tm + xe_pad)
        c = abi.AbcmbCosmo()
        p = params
        _set_inputs(c, p, k_pivot)
        c.has_ncdm = 1 if has_ncdm else 0
        c.species_off = int(species_off)
        if fluid_par is not None:
            c.fluid_par[:] = [float(v) for v in fluid_par]
        c.n_q = n_q if has_ncdm else 0
        c.xe_n = n_rec + xe_pad
        c.tm_n = n_tm + xe_pad
        c.tau_tab = _ptr(self.tau_tab).value
        c.xe_tab = _ptr(self.xe_tab).value
        c.tm_tab = _ptr(self.tm_tab).value
        if input_tau_reion:
            tau_re, z_re = p["tau_reion"], 0.0
        else:
            tau_re, z_re = -1.0, p["z_reion"]
        self._z_reion_in = z_re
        if lna_rec is not None and lna_transfer_start is not None:
            c.lna_rec, c.lna_transfer_start = lna_rec, lna_transfer_start
        sw = np.full(abi.REC_STAGES, -1, dtype=np.int32)
        if rec_switch is not None:
            sw[:] = rec_switch
        if reionization:
            rc = lib.abcmb_host_background(C.byref(c), tau_re, z_re, p["Delta_z_reion"], p["z_reion_He"],
                                           p["Delta_z_reion_He"], p["exp_reion"], p["Neff"], _ptr(sw))
        else:
            rc = lib.abcmb_host_recombination(C.byref(c), p["Neff"], _ptr(sw))
        self.reion = abi.AbcmbReion(tau_reion=tau_re, Delta_z_reion=p["Delta_z_reion"], z_reion_He=p["z_reion_He"],
                                    Delta_z_reion_He=p["Delta_z_reion_He"], exp_reion=p["exp_reion"])
        if rc != 0:
            raise RuntimeError(f"abcmb_host_background failed ({rc})")
        self.rec_switch = sw.copy()
        self.cosmo = c
        self.n_rec = n_rec

    # ------------------------------------------------------------ derivatives
    def tangents(self, dpar_in: np.ndarray, dtau_reion, dn_eff, reionization: bool = True) -> ProducerTangents:
        """Forward-mode derivatives of this background along the directions of
        ``dpar_in`` ([N_TANGENT_DIRS][COSMO_DOUBLES]: the derivatives of the
        parameter fields), ``dtau_reion`` and ``dn_eff`` ([N_TANGENT_DIRS]),
        with this background's grid choices held (abcmb_host_background_tangent)."""
        nd = abi.N_TANGENT_DIRS
        p = self.params
        c = abi.AbcmbCosmo.from_buffer_copy(self.cosmo)
        tau, xe, tm = np.zeros_like(self.tau_tab), np.zeros_like(self.xe_tab), np.zeros_like(self.tm_tab)
        c.tau_tab, c.xe_tab, c.tm_tab = _ptr(tau).value, _ptr(xe).value, _ptr(tm).value
        dpar = np.array(dpar_in, dtype=np.float64).reshape(nd, abi.COSMO_DOUBLES)
        dtr = np.ascontiguousarray(dtau_reion, dtype=np.float64).reshape(nd)
        dne = np.ascontiguousarray(dn_eff, dtype=np.float64).reshape(nd)
        dtau, dxe, dtm = np.zeros((nd, tau.size)), np.zeros((nd, xe.size)), np.zeros((nd, tm.size))
        sw = np.ascontiguousarray(self.rec_switch, dtype=np.int32)
        rc = self.lib.abcmb_host_background_tangent(
            C.byref(c), self.reion.tau_reion, _ptr(dtr), self._z_reion_in, p["Delta_z_reion"], p["z_reion_He"],
            p["Delta_z_reion_He"], p["exp_reion"], p["Neff"], _ptr(dne), _ptr(sw), 1 if reionization else 0,
            _ptr(dpar), _ptr(dtau), _ptr(dxe), _ptr(dtm))
        if rc != 0:
            raise RuntimeError(f"abcmb_host_background_tangent failed ({rc})")
        # the value parts repeat the producer's own arithmetic: they must equal
        # this background's tables bit for bit
        if not (np.array_equal(tau, self.tau_tab) and np.array_equal(xe, self.xe_tab)
                and np.array_equal(tm, self.tm_tab)):
            raise RuntimeError("abcmb_host_background_tangent: values differ from the background's")
        return ProducerTangents(dpar, dtau, dxe, dtm)

    def _tangent_call(self, fn, t: ProducerTangents, lna: np.ndarray, rows: int):
        lna = np.ascontiguousarray(np.atleast_1d(lna), dtype=np.float64)
        out = np.zeros((rows, lna.size))
        dout = np.zeros((abi.N_TANGENT_DIRS, rows, lna.size))
        rc = fn(C.byref(self.cosmo), _ptr(t.par), _ptr(t.tau), _ptr(t.xe), _ptr(t.tm), _ptr(lna), lna.size,
                _ptr(out), _ptr(dout))
        if rc != 0:
            raise RuntimeError("host tangent evaluation failed")
        return out, dout

    def los_vectors_tangent(self, t: ProducerTangents, lna: np.ndarray):
        """``los_vectors`` on ``lna`` ([6][n]) and its derivatives ([N_TANGENT_DIRS][6][n])."""
        return self._tangent_call(self.lib.abcmb_host_los_background_tangent, t, lna, 6)

    def eval_tangent(self, t: ProducerTangents, lna: np.ndarray):
        """``eval`` on ``lna`` ([7][n]: aH, tau, tau_c, cs2, x_e, T_m, R) and its
        derivatives ([N_TANGENT_DIRS][7][n])."""
        return self._tangent_call(self.lib.abcmb_host_eval_tangent, t, lna, 7)

    def shifted(self, params: dict, h: float, t: ProducerTangents, d: int) -> "Background":
        """This background moved by h along direction d of ``t``: every table and
        double field is value + h * derivative, the grid choices are this
        background's, the parameters are ``params``."""
        b = Background.__new__(Background)
        b.lib, b.params, b.n_rec, b.rec_switch = self.lib, params, self.n_rec, self.rec_switch.copy()
        b._z_reion_in = self._z_reion_in
        b.tau_tab = self.tau_tab + h * t.tau[d]
        b.xe_tab = self.xe_tab + h * t.xe[d]
        b.tm_tab = self.tm_tab + h * t.tm[d]
        c = abi.AbcmbCosmo.from_buffer_copy(self.cosmo)
        _doubles(c)[:] = _doubles(self.cosmo) + h * t.par[d]
        c.tau_tab, c.xe_tab, c.tm_tab = _ptr(b.tau_tab).value, _ptr(b.xe_tab).value, _ptr(b.tm_tab).value
        b.cosmo = c
        b.reion = abi.AbcmbReion.from_buffer_copy(self.reion)
        if self.reion.tau_reion > 0.0:
            b.reion.tau_reion = params["tau_reion"]
        return b

    # scalar accessors mirroring the reference's attribute names
    @property
    def tau0(self): return self.cosmo.tau0
    @property
    def lna_rec(self): return self.cosmo.lna_rec
    @property
    def lna_transfer_start(self): return self.cosmo.lna_transfer_start
    @property
    def z_reion(self): return self.cosmo.z_reion

    def los_vectors(self, lna: np.ndarray) -> np.ndarray:
        """[6, n]: tau, g, g', aH, exp(-kappa), aH' aH on ``lna``."""
        lna = np.ascontiguousarray(lna, dtype=np.float64)
        out = np.zeros((6, lna.size))
        rc = self.lib.abcmb_host_los_background(C.byref(self.cosmo), _ptr(lna), lna.size, _ptr(out))
        if rc != 0:
            raise RuntimeError("abcmb_host_los_background failed")
        return out

    def lna_start(self, k: np.ndarray, R_tc: float, R_large: float) -> np.ndarray:
        k = np.ascontiguousarray(k, dtype=np.float64)
        out = np.zeros(k.size)
        rc = self.lib.abcmb_host_lna_start(C.byref(self.cosmo), _ptr(k), k.size, R_tc, R_large, _ptr(out))
        if rc != 0:
            raise RuntimeError("abcmb_host_lna_start failed")
        return out

    def eval(self, lna: np.ndarray) -> dict:
        """aH, tau, tau_c, cs2, x_e, T_m, R at ``lna`` (host evaluation)."""
        lna = np.ascontiguousarray(np.atleast_1d(lna), dtype=np.float64)
        out = np.zeros((7, lna.size))
        self.lib.abcmb_host_eval(C.byref(self.cosmo), _ptr(lna), lna.size, _ptr(out))
        return dict(zip(("aH", "tau", "tau_c", "cs2", "xe", "Tm", "R"), out))

    def tau(self, lna):
        return self.eval(lna)["tau"]

    def R_ratio_lna(self, lna):
        """R = 3 rho_b / (4 rho_gamma) (abcmb/background.py:389-418)."""
        return 1.0 / self.eval(lna)["R"]

    def _drag(self):
        if getattr(self, "_drag_cache", None) is None:
            zd, rs = C.c_double(), C.c_double()
            if self.lib.abcmb_host_drag(C.byref(self.cosmo), C.byref(zd), C.byref(rs)) != 0:
                raise RuntimeError("abcmb_host_drag failed")
            self._drag_cache = (zd.value, rs.value)
        return self._drag_cache

    def z_d(self, params=None):
        """Baryon drag redshift, where the baryon optical depth reaches 1
        (abcmb/background.py:901-918)."""
        return self._drag()[0]

    def rs_d(self, params=None):
        """Sound horizon at the drag redshift in Mpc (abcmb/background.py:920-936)."""
        return self._drag()[1]
