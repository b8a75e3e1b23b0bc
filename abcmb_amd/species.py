"""Fluid species: the reference's plugin interface (abcmb/species.py:11-428).

A model's species are classes passed to ``Model(user_species=...)``
(abcmb/model_specs.py:89-116): the standard ones below, and user-defined
fluids that subclass ``Fluid``, ``StandardFluid``, ``BackgroundFluid`` or a
standard species, with the reference's method names and signatures:

    rho(lna, params), P(lna, params)              background, eV cm^-3
    y_ini(k, tau_ini, params)                     adiabatic initial conditions
    y_prime(k, lna, h', eta', y, args)            derivatives w.r.t. ln a,
                                                  args = (BG, params, species_list, species_dict)
    rho_delta / rho_plus_P_theta / rho_plus_P_sigma(lna, y, params)
    output_perturbations(lna, modes, args)

Write them with ``abcmb_amd.xp`` where the reference uses ``jax.numpy``
(``import abcmb_amd.xp as jnp``).  The model evaluates a user fluid's
methods once on symbolic inputs and compiles them into the GPU kernels
(``abcmb_amd.fluidgen``); the perturbation equations must therefore be linear
in the state, as the reference's are, and a user fluid's rows may couple to
other user fluids' rows and to the metric (h', eta'), not to the standard
species' rows.  The standard species' own rows are the kernels' built-in
code: their classes here carry the background densities and the names.
"""
from __future__ import annotations

from . import constants as cnst
from . import xp as jnp

RHO_UNIT = 3.0 * cnst.H0_over_h ** 2 / 8.0 / jnp.pi / cnst.G   # critical density per h^2, eV cm^-3


class Fluid:
    """Base class of every species (abcmb/species.py:11-249).

    ``first_idx``: index of the fluid's first row in the state vector, set by
    the model; ``num_equations``: its rows; ``name``: its key in
    ``species_dict`` and ``PT.species_perturbations``; ``is_matter``: counted
    in the matter density (P(k), omega_m).
    """
    first_idx = 0
    num_equations = 0
    name = ""
    is_matter = False

    def __init__(self, first_idx, specs):
        self.first_idx = first_idx

    def rho(self, lna, args):
        raise NotImplementedError("Fluid species must implement an energy density function.")

    def P(self, lna, args):
        raise NotImplementedError("Fluid species must implement a pressure function.")

    def w(self, lna, args):
        return self.P(lna, args) / self.rho(lna, args)

    def y_ini(self, k, tau_ini, args):
        raise NotImplementedError("Fluid species must implement the initial conditions of their perturbation modes.")

    def y_prime(self, k, lna, metric_h_prime, metric_eta_prime, y, args):
        raise NotImplementedError("Fluid species must implement a perturbation derivative function.")

    def rho_delta(self, lna, y, args):
        raise NotImplementedError("Fluid species must implement a perturbation derivative function.")

    def rho_plus_P_theta(self, lna, y, args):
        raise NotImplementedError("Fluid species must implement a perturbation derivative function.")

    def rho_plus_P_sigma(self, lna, y, args):
        raise NotImplementedError("Fluid species must implement a perturbation derivative function.")

    def output_perturbations(self, lna, modes, args):
        return {}


class StandardFluid(Fluid):
    """delta, theta, sigma in the fluid's first rows (abcmb/species.py:251-400)."""

    def get_delta(self, lna, y, args):
        return y[self.first_idx]

    def get_theta(self, lna, y, args):
        return jnp.where(self.num_equations > 1, y[self.first_idx + 1] if self.num_equations > 1 else 0.0, 0.0)

    def get_sigma(self, lna, y, args):
        return jnp.where(self.num_equations > 2, y[self.first_idx + 2] if self.num_equations > 2 else 0.0, 0.0)

    def rho_delta(self, lna, y, args):
        return self.rho(lna, args) * self.get_delta(lna, y, args)

    def rho_plus_P_theta(self, lna, y, args):
        return (self.rho(lna, args) + self.P(lna, args)) * self.get_theta(lna, y, args)

    def rho_plus_P_sigma(self, lna, y, args):
        return (self.rho(lna, args) + self.P(lna, args)) * self.get_sigma(lna, y, args)


class BackgroundFluid(Fluid):
    """A fluid without perturbations (abcmb/species.py:402-428)."""
    num_equations = 0

    def y_ini(self, k, tau_ini, args):
        return jnp.array([])

    def y_prime(self, k, lna, metric_h_prime, metric_eta_prime, y, args):
        return jnp.array([])

    def rho_delta(self, lna, y, args):
        return 0.0

    def rho_plus_P_theta(self, lna, y, args):
        return 0.0

    def rho_plus_P_sigma(self, lna, y, args):
        return 0.0


# ------------------------------------------------------------------ standard species
# Background densities of abcmb/species.py (DarkEnergy :432-490, ColdDarkMatter
# :492-610, MasslessNeutrino :612-760, MassiveNeutrino :762-1090, Baryon
# :1093-1291, Photon :1293-1456); their perturbation rows are the kernels'.

class DarkEnergy(BackgroundFluid):
    name = "DarkEnergy"

    def rho(self, lna, params):
        return params["omega_Lambda"] * RHO_UNIT

    def P(self, lna, params):
        return -self.rho(lna, params)


class ColdDarkMatter(StandardFluid):
    name = "ColdDarkMatter"
    num_equations = 1
    is_matter = True

    def rho(self, lna, params):
        return params["omega_cdm"] * RHO_UNIT / jnp.exp(lna) ** 3

    def P(self, lna, params):
        return 0.0

    def y_ini(self, k, tau_ini, params):
        return jnp.array([-(k * tau_ini) ** 2 / 4.0 * (1.0 - params["om"] * tau_ini / 5.0)])

    def y_prime(self, k, lna, metric_h_prime, metric_eta_prime, y, args):
        return jnp.array([-metric_h_prime / 2.0])

    def output_perturbations(self, lna, modes, args):
        return {"delta": modes[self.first_idx]}


class Baryon(StandardFluid):
    name = "Baryon"
    num_equations = 2
    is_matter = True

    def rho(self, lna, params):
        return params["omega_b"] * RHO_UNIT / jnp.exp(lna) ** 3

    def P(self, lna, params):
        return 0.0


class Photon(StandardFluid):
    name = "Photon"

    def __init__(self, first_idx, specs):
        super().__init__(first_idx, specs)
        self.num_equations = int(specs["l_max_g"]) + 1 + int(specs["l_max_pol_g"]) + 1

    def rho(self, lna, params):
        return jnp.pi ** 2 / 15.0 * params["TCMB0"] ** 4 / jnp.exp(lna) ** 4 / (cnst.c * cnst.hbar) ** 3

    def P(self, lna, params):
        return self.rho(lna, params) / 3.0


class MasslessNeutrino(StandardFluid):
    name = "MasslessNeutrino"

    def __init__(self, first_idx, specs):
        super().__init__(first_idx, specs)
        self.num_equations = int(specs["l_max_massless_nu"]) + 1

    def rho(self, lna, params):
        r = (params["N_nu_massless"] * 2.0 * 7.0 / 8.0 * jnp.pi ** 2 / 30.0 * params["T_nu_massless"] ** 4
             * params["TCMB0"] ** 4 / jnp.exp(lna) ** 4)
        return r / (cnst.c * cnst.hbar) ** 3

    def P(self, lna, params):
        return self.rho(lna, params) / 3.0


class MassiveNeutrino(Fluid):
    name = "MassiveNeutrino"
    is_matter = True

    def __init__(self, first_idx, specs):
        super().__init__(first_idx, specs)
        self.num_equations = int(specs.get("massive_nu_bins", 3)) * (int(specs["l_max_massive_nu"]) + 1)

    def _moments(self, lna, params):
        T = params["T_nu_massive"] * params["TCMB0"] / jnp.exp(lna)
        x = params["m_nu_massive"] / T
        ir = ip = 0.0
        for q, w in zip(cnst.Q_5P, cnst.W_5P):
            e = jnp.sqrt(q * q + x * x)
            ir = ir + w * (1.0 + jnp.exp(-q)) / (q * q) * e
            ip = ip + w * (1.0 + jnp.exp(-q)) / e
        pre = params["N_nu_massive"] * 4.0 * T ** 4 / jnp.pi ** 2 / cnst.hbar ** 3 / cnst.c ** 3
        return pre * ir, pre / 3.0 * ip

    def rho(self, lna, params):
        return self._moments(lna, params)[0]

    def P(self, lna, params):
        return self._moments(lna, params)[1]


LCDM_SPECIES = (DarkEnergy, ColdDarkMatter, Baryon, Photon, MasslessNeutrino)
STANDARD = (DarkEnergy, ColdDarkMatter, Baryon, Photon, MasslessNeutrino, MassiveNeutrino)


def is_standard(obj) -> bool:
    """True for the standard species themselves (not their subclasses)."""
    cls = obj if isinstance(obj, type) else type(obj)
    return cls in STANDARD
