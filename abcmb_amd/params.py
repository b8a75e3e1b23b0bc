"""Derived cosmological parameters.

Restates ``Model.add_derived_parameters`` (``abcmb/main.py:368-612``) for the
non-LINX paths (``bbn_type`` empty: Y_He is an input, default 0.245).  The
species sums run over the same species list, in the same order, as the
reference's loops (DarkEnergy, ColdDarkMatter, Baryon, Photon,
MasslessNeutrino, then MassiveNeutrino when it is a user species).
"""
from __future__ import annotations

import math

import numpy as np

from . import constants as cnst

_EXPECTED = {
    "h", "H0", "omega_cdm", "omega_b", "A_s", "n_s", "TCMB0", "tau_reion", "z_reion",
    "Delta_z_reion", "z_reion_He", "Delta_z_reion_He", "exp_reion", "omega_Lambda",
    "T_nu_massive", "N_nu_massive", "m_nu_massive", "N_nu_massless", "Neff",
    "T_nu_massless", "YHe", "omega_m", "R_b", "omega_r", "R_nu", "om",
}


def species_rho(name: str, lna: float, p: dict) -> float:
    """Energy density of one species (eV cm^-3), ``abcmb/species.py`` rho()."""
    a = math.exp(lna)
    if name == "DarkEnergy":
        return p["omega_Lambda"] * cnst.RHO_UNIT
    if name == "ColdDarkMatter":
        return p["omega_cdm"] * cnst.RHO_UNIT / a ** 3
    if name == "Baryon":
        return p["omega_b"] * cnst.RHO_UNIT / a ** 3
    if name == "Photon":
        return math.pi ** 2 / 15.0 * p["TCMB0"] ** 4 / a ** 4 / (cnst.c * cnst.hbar) ** 3
    if name == "MasslessNeutrino":
        r = p["N_nu_massless"] * 2.0 * 7.0 / 8.0 * math.pi ** 2 / 30.0 * p["T_nu_massless"] ** 4 * p["TCMB0"] ** 4 / a ** 4
        return r / (cnst.c * cnst.hbar) ** 3
    if name == "MassiveNeutrino":
        T = p["T_nu_massive"] * p["TCMB0"] / a
        x = p["m_nu_massive"] / T
        q = np.asarray(cnst.Q_5P)
        w = np.asarray(cnst.W_5P)
        integral = float(np.dot((1.0 + np.exp(-q)) / q ** 2 * np.sqrt(q ** 2 + x ** 2), w))
        return p["N_nu_massive"] * 4.0 * T ** 4 / math.pi ** 2 * integral / cnst.hbar ** 3 / cnst.c ** 3
    raise KeyError(name)


MATTER = {"ColdDarkMatter", "Baryon", "MassiveNeutrino"}


def _name(s) -> str:
    return s if isinstance(s, str) else s.name


def _rho(s, lna: float, p: dict) -> float:
    """A standard species by name, or a user fluid (abcmb_amd.species) through its own rho."""
    if isinstance(s, str):
        return species_rho(s, lna, p)
    return float(s.rho(lna, p))


def _matter(s) -> bool:
    return s in MATTER if isinstance(s, str) else bool(s.is_matter)


def add_derived_parameters(param_in: dict, species: tuple, specs: dict) -> dict:
    """Fill defaults and derived keys (``abcmb/main.py:368-612``).

    ``species``: the model's species in order, standard ones by name and user
    fluids as their instances (whose rho is called with numbers).

    Raises ``ValueError`` where the reference calls ``sys.exit()``
    (``abcmb/main.py:409-420``).
    """
    bbn = str(specs.get("bbn_type", "")).lower()
    if bbn in ("linx", "table"):
        raise NotImplementedError(
            f"bbn_type={specs['bbn_type']!r}: BBN (LINX / PArthENoPE table) is outside this "
            "library's scope; pass YHe explicitly with bbn_type=''")
    p = {k: float(v) if np.ndim(v) == 0 else v for k, v in param_in.items()}
    p["h"] = float(p.get("h", 0.6736))
    p["H0"] = p["h"] * cnst.H0_over_h
    p["omega_cdm"] = float(p.get("omega_cdm", 0.120))
    p["omega_b"] = float(p.get("omega_b", 0.02237))
    p["A_s"] = float(p.get("A_s", 2.1e-9))
    p["n_s"] = float(p.get("n_s", 0.9649))
    p["TCMB0"] = float(p.get("TCMB0", 2.34865418e-4))
    if specs["input_tau_reion"]:
        p["tau_reion"] = float(p.get("tau_reion", 0.0544))
    else:
        p["z_reion"] = float(p.get("z_reion", 7.67))
    p["Delta_z_reion"] = float(p.get("Delta_z_reion", 0.5))
    p["z_reion_He"] = float(p.get("z_reion_He", 3.5))
    p["Delta_z_reion_He"] = float(p.get("Delta_z_reion_He", 0.5))
    p["exp_reion"] = float(p.get("exp_reion", 1.5))
    p["omega_Lambda"] = 0.0
    p["T_nu_massive"] = float(p.get("T_nu_massive", 0.71611))
    p["N_nu_massive"] = float(p.get("N_nu_massive", 0))
    p["m_nu_massive"] = float(p.get("m_nu_massive", 0.06))

    input_N = param_in.get("N_nu_massless") is not None
    input_Neff = param_in.get("Neff") is not None
    if input_N and input_Neff:
        raise ValueError("You can only input one of N_nu_massless or Neff, got N_nu_massless="
                         f"{param_in['N_nu_massless']} and Neff={param_in['Neff']}.")
    if not input_N and not input_Neff:
        p["N_nu_massless"] = 3 - p["N_nu_massive"]
        input_N = True
    p["T_nu_massless"] = float(p.get("T_nu_massless", 0.71636856))

    lna_early = -23.0
    if input_N:
        rho_g = rho_nu = rho_extra = 0.0
        for s in species:
            rho = _rho(s, lna_early, p)
            if _name(s) == "Photon":
                rho_g += rho
            elif "neutrino" in _name(s).lower():
                rho_nu += rho
            else:
                rho_extra += rho
        fac = (8.0 / 7.0) * (11.0 / 4.0) ** (4.0 / 3.0)
        neff_raw = (rho_nu + rho_extra) / rho_g * fac
        rho_nu_early = 7 / 8 * (p["N_nu_massless"] + p["N_nu_massive"]) * p["T_nu_massless"] ** 4 * rho_g
        p["Neff"] = (rho_nu_early + rho_extra) / rho_g * fac
        p["N_nu_massless"] = p["N_nu_massless"] + p["Neff"] - neff_raw

    p["YHe"] = float(p.get("YHe", 0.245))

    if input_Neff:
        rho_g = rho_extra = 0.0
        for s in species:
            if _name(s) == "Photon":
                rho_g += _rho(s, lna_early, p)
            elif _name(s) != "MasslessNeutrino":
                rho_extra += _rho(s, lna_early, p)
        rho1nu = 7 / 8 * (4 / 11) ** (4 / 3) * rho_g
        p["N_nu_massless"] = (p["Neff"] - rho_extra / rho1nu) * ((4 / 11) ** (1 / 3) / p["T_nu_massless"]) ** 4

    rho_m = sum(_rho(s, 0.0, p) for s in species if _matter(s))
    p["omega_m"] = rho_m / cnst.RHO_UNIT
    p["R_b"] = p["omega_b"] / p["omega_m"]
    a_early = math.exp(-23.0)
    rho_r = rho_nu = 0.0
    for s in species:
        r = _rho(s, math.log(a_early), p)
        rho_r += r
        if "neutrino" in _name(s).lower():
            rho_nu += r
    p["omega_r"] = rho_r * a_early ** 4 / cnst.RHO_UNIT
    p["R_nu"] = rho_nu / rho_r
    p["om"] = p["omega_m"] / math.sqrt(p["omega_r"]) * cnst.H0_over_h / cnst.c_Mpc_over_s
    p["omega_Lambda"] = p["h"] ** 2 - p["omega_r"] - p["omega_m"]
    for key, val in param_in.items():
        if key not in _EXPECTED:
            p[key] = val
    return p
