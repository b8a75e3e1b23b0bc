"""GPU discrete adjoint of the perturbation ODE (abcmb_pert_adjoint) through the C ABI.

The adjoint is pinned to the tangent-linear kernel, which the oracle pins
(test_gpu_tangent.py): for any table cotangent lam and any tangent seed s of
the host inputs (cosmology fields and the tau, x_e, T_m tables),

    <lam, J s> = <J^T lam, s>,

J the derivative of the perturbation tables with respect to those inputs at
the launch's step sequence.  The tangent kernel gives the left side along
five seed directions, the adjoint the right side for all of them at once.
Both solve their stage equations to 1e-6 (relative, per stage), so the two
sides agree to about that fraction of sum |lam * J s|; the bound is 2e-5.
k-modes whose tangent-kernel step sequence differs from the primal's (a
rounding-level accept/reject flip between the two kernels) get lam = 0.
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu

# the AbcmbCosmo fields the perturbation equations read (bg_eval, fill_scalars,
# ic_row, all_fields)
ODE_FIELDS = ("omega_b", "omega_cdm", "omega_Lambda", "TCMB0", "YHe", "N_ur", "T_ur", "N_ncdm", "T_ncdm",
              "m_ncdm", "xe_last", "tm_last", "om", "R_nu", "R_b")


def _say(msg):
    print(msg, flush=True)


def _random_seeds(plan, rng):
    """Five smooth random tangent directions of the host inputs."""
    from abcmb_amd import abi
    from abcmb_amd.model import TangentSeeds, cosmo_doubles
    nd = abi.N_TANGENT_DIRS
    names = [f[0] for f in abi.AbcmbCosmo._fields_][:abi.COSMO_DOUBLES]
    vals = cosmo_doubles(plan.bg.cosmo)
    par = np.zeros((nd, abi.COSMO_DOUBLES))
    for f in ODE_FIELDS:
        i = names.index(f)
        par[:, i] = rng.normal(size=nd) * 1e-2 * (abs(vals[i]) + 1e-30)

    def smooth(tab):
        x = np.linspace(0.0, 1.0, tab.size)
        c = rng.normal(size=(nd, 3)) * 1e-2
        return tab[None] * (c[:, :1] + c[:, 1:2] * x[None] + c[:, 2:3] * np.sin(7.0 * x)[None])

    return TangentSeeds(par, smooth(plan.bg.tau_tab), smooth(plan.bg.xe_tab), smooth(plan.bg.tm_tab))


def _dot_identity(m, plan, seeds, rng, label):
    import torch
    outs, dpts = m.execute_tangent([plan], [seeds])
    dpt = dpts[0]
    pt = outs[0].PT.table
    prim = m.execute([plan])[0]
    same = torch.as_tensor(outs[0].steps == prim.steps, device=pt.device)
    # a table cotangent with every field weighted by its size
    rms = pt.pow(2).mean(dim=(1, 2)).sqrt()
    w = torch.where(rms > 0, 1.0 / rms, torch.zeros_like(rms))
    lam = torch.randn(pt.shape, generator=torch.Generator(device="cpu").manual_seed(int(rng.integers(1 << 30))),
                      dtype=torch.float64).to(pt.device) * w[:, None, None]
    lam = lam * same[None, None, :].to(lam.dtype)
    _say(f"case: {label}, {int(same.sum())} of {same.numel()} k-modes with the primal's step sequence")
    cot = m.ode_adjoint([plan], [lam])[0]
    assert np.array_equal(cot["steps"], prim.steps), "the adjoint's forward pass left the primal's step sequence"
    dev = pt.device
    sp = torch.as_tensor(seeds.par, device=dev)
    st = torch.as_tensor(seeds.tau, device=dev)
    sx = torch.as_tensor(seeds.xe, device=dev)
    sm = torch.as_tensor(seeds.tm, device=dev)
    for d in range(sp.shape[0]):
        lhs = float(torch.sum(lam * dpt[d]))
        rhs = float(cot["par"] @ sp[d] + cot["tau"] @ st[d] + cot["xe"] @ sx[d] + cot["tm"] @ sm[d])
        scale = float(torch.sum(torch.abs(lam * dpt[d])))
        _say(f"  direction {d}: <lam, J s> = {lhs:.12e}, <J^T lam, s> = {rhs:.12e}, "
             f"difference / sum|lam J s| = {abs(lhs - rhs) / scale:.2e}")
        assert abs(lhs - rhs) <= 2e-5 * scale


@pytest.mark.parametrize("massive", [False, True])
def test_adjoint_is_the_transpose_of_the_tangent(massive):
    from abcmb_amd.model import Model, MassiveNeutrino
    kk = np.geomspace(1e-4, 0.3, 12)
    m = Model(user_species=(MassiveNeutrino,) if massive else (), l_max=500, k_axis_perturbations=kk)
    plan = m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))
    rng = np.random.default_rng(11 + massive)
    _dot_identity(m, plan, _random_seeds(plan, rng), rng, f"random input directions, massive={massive}")


def test_adjoint_parameter_directions():
    """The same identity along the five parameter directions the gradients use
    (host seeds by central differences of the producer)."""
    from abcmb_amd.model import Model, MassiveNeutrino, tangent_seeds
    kk = np.geomspace(1e-4, 0.3, 12)
    m = Model(user_species=(MassiveNeutrino,), l_max=500, k_axis_perturbations=kk)
    th = LK.FIDUCIAL.copy()
    base = m.prepare(m.derived(LK.theta_to_params(th)))
    steps = LK.FD_REL_STEP * np.abs(th)
    plus, minus = [], []
    for p in LK.ODE_DIRS:
        for sgn, lst in ((1, plus), (-1, minus)):
            t = th.copy()
            t[p] += sgn * steps[p]
            lst.append(m.prepare(m.derived(LK.theta_to_params(t)), freeze=base))
    seeds = tangent_seeds(plus, minus, [steps[p] for p in LK.ODE_DIRS])
    _dot_identity(m, base, seeds, np.random.default_rng(5), "parameter directions, massive nu")
