"""Run options ("specs"), state-vector layout and the k-grids.

Mirrors ``abcmb/model_specs.py``: the same spec keys and defaults
(``load_specs`` :7-82), the same species ordering in the ODE state
(``populate_species`` :84-118) and the same k-grid construction loops
(``get_k_axis_perturbations`` :120-175, ``get_k_axis_transfer`` :177-193).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_DEFAULTS = dict(
    use_LCDM_species=True,
    input_tau_reion=True,
    l_min=2,
    l_max=2500,
    lensing=False,
    k_max=0.5,
    bbn_type="",
    linx_reaction_net="key_PRIMAT_2023",
    l_max_g=12,
    l_max_pol_g=10,
    l_max_massless_nu=17,
    l_max_massive_nu=17,
    k_step_sub=5.0e-2,
    k_step_super=2.0e-3,
    k_step_transition=2.0e-1,
    k_step_super_reduction=1.0e-1,
    k_min_tau0=1.0e-1,
    k_max_tau0_over_l_max=1.8,
    H0_fid=2.255560e-04,
    tau0_fid=1.418668e04,
    rs_rec_fid=1.446279e02,
    k_transfer_linstep=4.5e-1,
    k_transfer_logstep=170.0,
    tau_rec_fid=281.040565,
    k_pivot=0.05,
    R_tc=0.0015,
    R_large=0.07,
    max_steps_PE=2048,
    k_split_PE=0.01,
    rtol_small_k_PE=1.0e-5,
    rtol_large_k_PE=1.0e-4,
    atol_small_k_PE=1.0e-10,
    atol_large_k_PE=1.0e-6,
    pcoeff_PE=0.25,
    icoeff_PE=0.8,
    dcoeff_PE=0.0,
    scale_sw=1,
    scale_isw=1,
    scale_dop=1,
    scale_pol=1,
)


def load_specs(input_specs: dict) -> dict:
    """Fill defaults; unknown keys are kept (``abcmb/model_specs.py:7-82``).

    Extensions of this build (all optional, absent from the reference):
    ``k_axis_perturbations`` (explicit PT k-grid, Mpc^-1), ``massive_nu_bins``
    (3 = reference quadrature, 5 = the 5-point quadrature), ``n_lna_save``
    (500 in the reference, ``abcmb/perturbations.py:100``).
    """
    specs = {}
    for key, val in _DEFAULTS.items():
        specs[key] = input_specs.get(key, val)
    for key, val in input_specs.items():
        if key not in specs:
            specs[key] = val
    specs.setdefault("massive_nu_bins", 3)
    specs.setdefault("n_lna_save", 500)
    return specs


@dataclass(frozen=True)
class StateLayout:
    """Indices of the ODE state vector.

    Order is the reference's species order: metric eta, then DarkEnergy (no
    equations), ColdDarkMatter, Baryon, Photon (F then G), MasslessNeutrino,
    then an optional MassiveNeutrino user species
    (``abcmb/model_specs.py:84-118``; ``abcmb/perturbations.py:194-196``).
    """
    l_g: int
    l_pol: int
    l_ur: int
    l_ncdm: int
    n_q: int            # 0 when no massive-neutrino species is present

    @property
    def i_eta(self): return 0
    @property
    def i_cdm(self): return 1
    @property
    def i_b(self): return 2
    @property
    def i_g(self): return 4
    @property
    def i_pol(self): return 4 + self.l_g + 1
    @property
    def i_ur(self): return self.i_pol + self.l_pol + 1
    @property
    def i_ncdm(self): return self.i_ur + self.l_ur + 1
    @property
    def n(self):
        return self.i_ncdm + self.n_q * (self.l_ncdm + 1)


def state_layout(specs: dict, massive: bool) -> StateLayout:
    n_q = int(specs.get("massive_nu_bins", 3)) if massive else 0
    return StateLayout(int(specs["l_max_g"]), int(specs["l_max_pol_g"]),
                       int(specs["l_max_massless_nu"]), int(specs["l_max_massive_nu"]), n_q)


def get_k_axis_perturbations(specs: dict):
    """PT k-grid and P(k) output grid (``abcmb/model_specs.py:120-175``).

    Sets ``specs['k_min']`` and ``specs['k_max_cmb']`` as the reference does.
    """
    ks = np.zeros(2000)
    H0_fid = specs["H0_fid"]
    tau0_fid = specs["tau0_fid"]
    k_rec_fid = 2.0 * math.pi / specs["rs_rec_fid"]
    k_min = specs["k_min_tau0"] / tau0_fid
    k_max = specs["k_max_tau0_over_l_max"] / tau0_fid * specs["l_max"]
    k = k_min
    ks[0] = k
    i = 0
    while k < k_max:
        step = (specs["k_step_super"]
                + 0.5 * (math.tanh((k - k_rec_fid) / k_rec_fid / specs["k_step_transition"]) + 1.0)
                * (specs["k_step_sub"] - specs["k_step_super"])) * k_rec_fid
        scale2 = H0_fid ** 2
        step *= (k ** 2 / scale2 + 1.0) / (k ** 2 / scale2 + 1.0 / specs["k_step_super_reduction"])
        k += step
        i += 1
        ks[i] = k
    specs["k_min"] = k_min
    specs["k_max_cmb"] = k
    if specs["lensing"]:
        k_max = k + 0.3
        while k < k_max:
            k += 0.005
            i += 1
            ks[i] = k
    if k < specs["k_max"]:
        k_max = specs["k_max"]
        while k < k_max:
            k += 0.005
            i += 1
            ks[i] = k
    ks = ks[np.where(ks > 0)]
    k_pk = ks[np.where(ks <= specs["k_max"])]
    return ks, k_pk


def get_k_axis_transfer(specs: dict) -> np.ndarray:
    """Transfer (LOS) k-grid, shared by all cosmologies (``abcmb/model_specs.py:177-193``)."""
    ks = np.zeros(8000)
    k_period = 2 * math.pi / (specs["tau0_fid"] - specs["tau_rec_fid"])
    k = specs["k_min"]
    ks[0] = k
    i = 0
    while k < specs["k_max_cmb"]:
        k = k + k_period * specs["k_transfer_linstep"] * k / (k + specs["k_transfer_linstep"] / specs["k_transfer_logstep"])
        i += 1
        ks[i] = k
    return ks[np.where(ks > 0)]
