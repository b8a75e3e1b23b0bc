"""The autograd front end (``abcmb_amd.autograd.lensed_cl``) on the GPU.

mode="jacobian":
  - its forward values are the spectra ``tangent_gradients`` observes, bit for
    bit;
  - ``LensedClLikelihood.loglike_cl`` applied to them and back-propagated with
    ``backward()`` gives ``tangent_gradients``' likelihood gradients to 1e-6 of
    the norm (the Gaussian likelihood is quadratic in the spectra, so the two
    differ only by the likelihood's curvature times the O(h^2) offset of the
    mean of the moved lensed spectra from theta's; measured 4e-9);
  - a second loss through the same graph (a weighted sum of TT) gets its own
    gradient from the same Jacobian, equal to the Jacobian contracted by hand.

mode="adjoint" (the default, reverse mode in the backward pass):
  - the likelihood's gradient equals ``adjoint_gradients`` to rounding and
    ``tangent_gradients`` to 1e-4 of the norm (two linearisations of the same
    discretisation, measured 1e-6 to 5e-5);
  - the backward leaves the host inputs' cotangents on the function node, and
    contracting them with the host seeds by hand gives the same gradient.
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_autograd_matches_tangent_gradients():
    import torch
    from abcmb_amd.autograd import lensed_cl
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=1000)
    thetas = LK.latin_hypercube(3, seed=61)
    kg = [k[::4] for k in LK.uneven_k_grids(3, seed=61, hi=0.3)]
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=3, seed=1)
    _say("case: 3 cosmologies, tangent_gradients vs the autograd front end")
    cl_ref = []
    L, g = LK.tangent_gradients(m, like, thetas, k_grids=kg, cl_out=cl_ref)
    cl_ref = torch.cat(cl_ref)
    theta = torch.tensor(thetas, dtype=torch.float64, requires_grad=True)
    cl = lensed_cl(m, theta, k_grids=kg, mode="jacobian")
    assert torch.equal(cl.detach().to(cl_ref.device), cl_ref)
    loss = like.loglike_cl(cl.cuda(), [0, 1, 2])
    assert np.max(np.abs(loss.detach().cpu().numpy() / L - 1)) < 1e-13
    loss.sum().backward()
    ga = theta.grad.numpy()
    scale = np.linalg.norm(g * thetas, axis=1, keepdims=True) / np.abs(thetas)
    err = np.max(np.abs(ga - g) / scale)
    _say(f"autograd vs tangent_gradients: {err:.1e} of the norm")
    assert err < 1e-6
    # another loss through the same function
    theta2 = torch.tensor(thetas, dtype=torch.float64, requires_grad=True)
    w = torch.linspace(0.0, 1.0, m.ells.size, dtype=torch.float64, device="cuda")
    cl2 = lensed_cl(m, theta2, k_grids=kg, mode="jacobian")
    (cl2[:, 0].cuda() * w).sum().backward()
    jac = []
    LK.tangent_gradients(m, like, thetas, k_grids=kg, jac_out=jac)
    jac = torch.cat(jac)
    by_hand = (jac[:, :, 0] * w).sum(dim=-1).cpu().numpy()
    assert np.max(np.abs(theta2.grad.numpy() - by_hand) / np.max(np.abs(by_hand), axis=1, keepdims=True)) < 1e-12


def test_autograd_adjoint_mode():
    import torch
    from abcmb_amd.autograd import lensed_cl
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=1000)
    thetas = LK.latin_hypercube(3, seed=61)
    kg = [k[::4] for k in LK.uneven_k_grids(3, seed=61, hi=0.3)]
    fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
    like = LK.LensedClLikelihood(m, [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()],
                                 n_mocks=3, seed=1)
    _say("case: 3 cosmologies, the adjoint-mode autograd front end")
    La, ga = LK.adjoint_gradients(m, like, thetas, k_grids=kg)
    Lt, gt = LK.tangent_gradients(m, like, thetas, k_grids=kg)
    theta = torch.tensor(thetas, dtype=torch.float64, requires_grad=True)
    cl = lensed_cl(m, theta, k_grids=kg)
    loss = like.loglike_cl(cl.cuda(), [0, 1, 2])
    assert np.max(np.abs(loss.detach().cpu().numpy() / La - 1)) < 1e-13
    loss.sum().backward()
    g = theta.grad.numpy()
    scale = np.linalg.norm(gt * thetas, axis=1, keepdims=True) / np.abs(thetas)
    e_adj = np.max(np.abs(g - ga) / scale)
    e_tan = np.max(np.abs(g - gt) / scale)
    _say(f"autograd (adjoint) vs adjoint_gradients {e_adj:.1e}, vs tangent_gradients {e_tan:.1e} of the norm")
    assert e_adj < 1e-12
    assert e_tan < 1e-4
    cots = cl.grad_fn.host_cotangents
    assert len(cots) == 3 and all(np.all(np.isfinite(c["tau"])) for c in cots)
    # the same omega_b gradient by contracting the exposed cotangents by hand: the
    # host seeds, the LOS background vectors and C_l^phiphi's background part
    steps = LK.FD_REL_STEP * np.abs(thetas)
    state, _ = LK.adjoint_forward(m, thetas, kg)
    pm, seeds = LK._seed_inputs(m, thetas, state["plans"], kg, steps)
    clpp_d = LK._clpp_directional(m, thetas, steps, state["plans"], pm, state["run"]["pts"],
                                  state["run"]["kdev"]).cpu().numpy()
    d = LK.ODE_DIRS.index(0)
    for i in range(3):
        c = cots[i]
        dbgv = (pm[i][2 * d].bgv - pm[i][2 * d + 1].bgv) / (2.0 * steps[i, 0])
        by_hand = (c["par"] @ seeds[i].par[d] + c["tau"] @ seeds[i].tau[d] + c["xe"] @ seeds[i].xe[d] +
                   c["tm"] @ seeds[i].tm[d] + np.sum(c["bgv"] * dbgv) + c["clpp"] @ clpp_d[i, 0])
        assert abs(by_hand - g[i, 0]) <= 1e-10 * scale[i, 0]
