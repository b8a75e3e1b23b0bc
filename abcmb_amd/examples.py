"""Example user-defined fluids (the models of the reference's fluid tutorial,
example_notebooks/ABCMB_Fluids.ipynb), written against abcmb_amd.species.

  * ``W0WaDarkEnergy``: dark energy with w(a) = w0 + (1 - a) wa (CPL), a
    background fluid replacing the cosmological constant
    (``Model(user_species=(ColdDarkMatter, Baryon, Photon, MasslessNeutrino,
    W0WaDarkEnergy), use_LCDM_species=False)``, parameters ``w0DE``, ``waDE``);
  * ``SelfInteractingDR``: tightly self-coupled dark radiation with density
    and velocity perturbations only, its density N_idr times one neutrino
    species' (parameter ``N_idr``);
  * ``InteractingDM`` / ``InteractingDR``: dark matter with a velocity that
    scatters on a dark-radiation fluid at the rate Gamma_0_nadm / a (per unit
    conformal time, in Mpc^-1), momentum conserved by R = 3 rho_dm / (4 rho_dr);
    together they replace cold dark matter (use_LCDM_species=False with
    DarkEnergy, Baryon, Photon, MasslessNeutrino).
"""
from __future__ import annotations

from . import constants as cnst
from . import xp as jnp
from .species import BackgroundFluid, ColdDarkMatter, RHO_UNIT, StandardFluid


class W0WaDarkEnergy(BackgroundFluid):
    name = "W0WaDarkEnergy"

    def rho(self, lna, params):
        # d ln rho / d ln a = -3 (1 + w(a)) integrated from a = 1
        a = jnp.exp(lna)
        w0, wa = params["w0DE"], params["waDE"]
        return params["omega_Lambda"] * RHO_UNIT * a ** (-3.0 * (1.0 + w0 + wa)) * jnp.exp(3.0 * wa * (a - 1.0))

    def P(self, lna, params):
        a = jnp.exp(lna)
        return (params["w0DE"] + (1.0 - a) * params["waDE"]) * self.rho(lna, params)


def _one_neutrino(lna, params):
    """rho of one massless neutrino species at the standard temperature ratio."""
    a = jnp.exp(lna)
    rho_g = jnp.pi ** 2 / 15.0 * params["TCMB0"] ** 4 / a ** 4 / (cnst.hbar * cnst.c) ** 3
    return 7.0 / 8.0 * (4.0 / 11.0) ** (4.0 / 3.0) * rho_g


class SelfInteractingDR(StandardFluid):
    name = "SelfInteractingDR"

    def __init__(self, first_idx, specs):
        super().__init__(first_idx, specs)
        self.num_equations = 2

    def rho(self, lna, params):
        return params["N_idr"] * _one_neutrino(lna, params)

    def P(self, lna, params):
        return self.rho(lna, params) / 3.0

    def y_ini(self, k, tau, params):
        # adiabatic: as massless neutrinos' delta and theta
        Rn, om = params["R_nu"], params["om"]
        delta = -(k * tau) ** 2 / 3.0 * (1.0 - om * tau / 5.0)
        theta = (-k * (k * tau) ** 3 / 36.0 / (4.0 * Rn + 15.0)
                 * (4.0 * Rn + 23.0 - 3.0 * (8.0 * Rn ** 2 + 50.0 * Rn + 275.0) / 20.0 / (2.0 * Rn + 15.0) * tau * om))
        return jnp.array([delta, theta])

    def _derivs(self, k, aH, hp, y, coupling):
        delta, theta = y[self.first_idx], y[self.first_idx + 1]
        d_delta = -4.0 / 3.0 * theta / aH - 2.0 / 3.0 * hp
        d_theta = k ** 2 / 4.0 * delta / aH + coupling
        return jnp.array([d_delta, d_theta])

    def y_prime(self, k, lna, hp, ep, y, args):
        BG, params, _, _ = args
        return self._derivs(k, BG.aH(lna, params), hp, y, 0.0)

    def output_perturbations(self, lna, modes, args):
        return {"delta": modes[self.first_idx], "theta": modes[self.first_idx + 1]}


class InteractingDM(ColdDarkMatter):
    name = "InteractingDM"
    num_equations = 2

    def y_ini(self, k, tau, params):
        om, Rb, Rn = params["om"], params["R_b"], params["R_nu"]
        delta = -(k * tau) ** 2 / 4.0 * (1.0 - om * tau / 5.0)
        theta = -k ** 4 * tau ** 3 / 36.0 * (1.0 - 3.0 * (1.0 + 5.0 * Rb - Rn) / 20.0 / (1.0 - Rn) * om * tau)
        return jnp.array([delta, theta])

    def y_prime(self, k, lna, hp, ep, y, args):
        BG, params, species, index = args
        dr = species[index["InteractingDR"]]
        aH = BG.aH(lna, params)
        rate = params["Gamma_0_nadm"] / jnp.exp(lna) / aH
        delta, theta = y[self.first_idx], y[self.first_idx + 1]
        theta_dr = y[dr.first_idx + 1]
        return jnp.array([-theta / aH - hp / 2.0, -theta + rate * (theta_dr - theta)])

    def output_perturbations(self, lna, modes, args):
        return {"delta": modes[self.first_idx], "theta": modes[self.first_idx + 1]}


class InteractingDR(SelfInteractingDR):
    name = "InteractingDR"

    def y_prime(self, k, lna, hp, ep, y, args):
        BG, params, species, index = args
        dm = species[index["InteractingDM"]]
        aH = BG.aH(lna, params)
        R = 3.0 * dm.rho(lna, params) / (4.0 * self.rho(lna, params))
        rate = params["Gamma_0_nadm"] / jnp.exp(lna) / aH
        theta, theta_dm = y[self.first_idx + 1], y[dm.first_idx + 1]
        return self._derivs(k, aH, hp, y, R * rate * (theta_dm - theta))


# the tutorial's models as Model keyword arguments (with their parameters);
# __graft_entry__.build() compiles their kernels ahead of the GPU tests
MODELS = {
    "sidr": dict(user_species=(SelfInteractingDR,)),
    "nadm": dict(user_species=("DarkEnergy", InteractingDM, InteractingDR, "Baryon", "Photon", "MasslessNeutrino"),
                 use_LCDM_species=False),
    "w0wa": dict(user_species=("ColdDarkMatter", "Baryon", "Photon", "MasslessNeutrino", W0WaDarkEnergy),
                 use_LCDM_species=False),
}
PARAMS = {"sidr": dict(N_idr=0.3), "nadm": dict(N_idr=0.3, Gamma_0_nadm=4e-7), "w0wa": dict(w0DE=-0.8, waDE=-0.6)}


def model_kwargs(name: str) -> dict:
    """MODELS[name] with the standard species' names resolved to their classes."""
    from . import species as SP
    std = {c.name: c for c in SP.STANDARD}
    kw = dict(MODELS[name])
    kw["user_species"] = tuple(std[c] if isinstance(c, str) else c for c in kw["user_species"])
    return kw
