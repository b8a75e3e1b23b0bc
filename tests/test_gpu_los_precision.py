"""Reduced-precision line-of-sight variants in the gradient path (configuration 5's
FP32 and BF16 LOS contractions), through ``tangent_gradients`` on three
Latin-hypercube cosmologies lensed to l = 2500: the observed spectra and the
likelihood gradients of the FP32 and BF16-operand tangent-linear LOS against the
FP64 one (same tables, same lensing).  Measured: FP32 C_l TT 9.4e-6, TE 2.8e-7,
EE 1.1e-5 (relative per l; TE of its maximum) and the gradient 1.0e-3 of the norm
(the likelihood gradient sums C_l derivatives with large weights of both signs,
which amplifies the FP32 phase error of the Bessel functions at large k chi);
BF16 operands C_l TT 1.1e-2, TE 1.0e-2, EE 8.6e-4 and the gradient 1.0e-2.
Bounds: FP32 spectra within the 1e-4 contract and the gradient within 3e-3;
BF16 within 5e-2 and 1e-1.
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_reduced_precision_los_gradients():
    import torch
    from abcmb_amd.model import Model, MassiveNeutrino
    thetas = LK.latin_hypercube(3, seed=71)
    kg = [k[::2] for k in LK.uneven_k_grids(3, seed=71)]
    res = {}
    for prec in ("fp64", "fp32", "bf16"):
        m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096, los_precision=prec)
        if prec == "fp64":
            fid = m.execute([m.prepare(m.derived(LK.theta_to_params(LK.FIDUCIAL)))])[0]
            obs = [fid.ClTT.cpu().numpy(), fid.ClTE.cpu().numpy(), fid.ClEE.cpu().numpy()]
        like = LK.LensedClLikelihood(m, obs, n_mocks=3, seed=1)
        _say(f"case: tangent gradients of 3 cosmologies, {prec} LOS")
        cl = []
        _, g = LK.tangent_gradients(m, like, thetas, k_grids=kg, cl_out=cl)
        res[prec] = (torch.cat(cl).cpu().numpy(), g)
    cl64, g64 = res["fp64"]
    scale = np.linalg.norm(g64 * thetas, axis=1, keepdims=True) / np.abs(thetas)
    out = {}
    for prec in ("fp32", "bf16"):
        cl, g = res[prec]
        e = [float(np.max(np.abs(cl[:, x] - cl64[:, x]) / (np.max(np.abs(cl64[:, x]), axis=1, keepdims=True)
                                                            if x == 1 else np.abs(cl64[:, x])))) for x in range(3)]
        ge = float(np.max(np.abs(g - g64) / scale))
        out[prec] = (e, ge)
        _say(f"{prec} LOS vs FP64: C_l TT {e[0]:.1e} TE {e[1]:.1e} EE {e[2]:.1e}; gradient {ge:.1e} of the norm")
        assert np.all(np.isfinite(cl)) and np.all(np.isfinite(g))
    assert max(out["fp32"][0]) < 1e-4 and out["fp32"][1] < 3e-3
    assert max(out["bf16"][0]) < 5e-2 and out["bf16"][1] < 1e-1
