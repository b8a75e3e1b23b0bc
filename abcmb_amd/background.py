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
        c.h, c.omega_b, c.omega_cdm, c.omega_Lambda = p["h"], p["omega_b"], p["omega_cdm"], p["omega_Lambda"]
        c.A_s, c.n_s, c.k_pivot, c.TCMB0, c.YHe = p["A_s"], p["n_s"], k_pivot, p["TCMB0"], p["YHe"]
        c.N_ur, c.T_ur = p["N_nu_massless"], p["T_nu_massless"]
        c.N_ncdm, c.T_ncdm, c.m_ncdm = p["N_nu_massive"], p["T_nu_massive"], p["m_nu_massive"]
        c.R_nu, c.R_b, c.om, c.omega_r, c.omega_m = p["R_nu"], p["R_b"], p["om"], p["omega_r"], p["omega_m"]
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
