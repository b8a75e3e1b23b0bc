"""The autograd front end (``abcmb_amd.autograd.lensed_cl``) on the GPU:
  - its forward values are the spectra ``tangent_gradients`` observes, bit for
    bit;
  - ``LensedClLikelihood.loglike_cl`` applied to them and back-propagated with
    ``backward()`` gives ``tangent_gradients``' likelihood gradients to 1e-6 of
    the norm (the Gaussian likelihood is quadratic in the spectra, so the two
    differ only by the likelihood's curvature times the O(h^2) offset of the
    mean of the moved lensed spectra from theta's; measured 4e-9);
  - a second loss through the same graph (a weighted sum of TT) gets its own
    gradient from the same Jacobian, equal to the Jacobian contracted by hand.
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
    cl = lensed_cl(m, theta, k_grids=kg)
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
    cl2 = lensed_cl(m, theta2, k_grids=kg)
    (cl2[:, 0].cuda() * w).sum().backward()
    jac = []
    LK.tangent_gradients(m, like, thetas, k_grids=kg, jac_out=jac)
    jac = torch.cat(jac)
    by_hand = (jac[:, :, 0] * w).sum(dim=-1).cpu().numpy()
    assert np.max(np.abs(theta2.grad.numpy() - by_hand) / np.max(np.abs(by_hand), axis=1, keepdims=True)) < 1e-12
