"""Batching machinery on the GPU, through the C ABI:
  - the persistent ODE kernel (a grid of resident workgroups taking items
    from a device work queue) gives the one-workgroup-per-item launch's
    tables, step counts and chord-iteration counts bit for bit, for the
    value-only and the tangent-linear kernel, with a grid both larger and
    much smaller than the item count;
  - the downstream captured as a HIP graph and replayed for other inputs
    (with eager work and synchronisations between the replays) gives the
    eager results bit for bit on every replay.
"""
import numpy as np
import pytest

from abcmb_amd import likelihood as LK

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _tune(key, value):
    from abcmb_amd import abi
    abi.check(abi.hip_lib().abcmb_tune_set(key, value), "abcmb_tune_set")


@pytest.mark.parametrize("tangent", [False, True])
def test_persistent_queue_matches_per_item_launch(tangent):
    import torch
    from abcmb_amd import abi
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=False, l_max=1000, max_steps_PE=4096)
    thetas = LK.latin_hypercube(3, seed=21)
    kg = [np.geomspace(1e-5, 0.3, n) for n in (40, 57, 33)]
    steps = LK.FD_REL_STEP * np.abs(thetas)
    plans, pm, seeds = LK._tangent_inputs(m, thetas, kg, steps)
    runs = {}
    try:
        for sched in (0, 100, 3):
            _tune(abi.TUNE_PERT_PERSISTENT, sched)
            _say(f"case: ODE launch, {'tangent' if tangent else 'value'} kernel, schedule {sched}")
            h = m._launch_ode(plans, tangents=seeds if tangent else None)
            r = m._finish_ode(h)
            runs[sched] = (r, h["iters"].cpu().numpy())
    finally:
        _tune(abi.TUNE_PERT_PERSISTENT, 0)
    base, it0 = runs[0]
    for sched in (100, 3):
        r, it = runs[sched]
        assert np.array_equal(it, it0)
        for c in range(3):
            assert np.array_equal(r["steps"][c], base["steps"][c])
            assert torch.equal(r["pts"][c], base["pts"][c])
            if tangent:
                assert torch.equal(r["dpts"][c], base["dpts"][c])


def test_downstream_graph_replays_match_eager():
    import torch
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=1000, max_steps_PE=4096)
    sets = []
    for seed in (31, 32):
        thetas = LK.latin_hypercube(3, seed=seed)
        kg = LK.uneven_k_grids(3, seed=seed)
        kg = [k[::8] for k in kg]          # 38-100 modes each
        plans = m.prepare_many([m.derived(LK.theta_to_params(t)) for t in thetas], k_list=kg)
        run = m._run_ode(plans)
        eager = m._downstream(plans, run["pts"], run["kdev"])
        sets.append((plans, run, {key: v.clone() for key, v in eager.items() if v is not None}))
    nkmax = max(int(k.numel()) for _, run, _ in sets for k in run["kdev"])
    _say(f"case: downstream graph of 3 cosmologies (nk <= {nkmax}), replays A, B, A")
    g = m.downstream_graph(3, nkmax)
    for rep, idx in enumerate((0, 1, 0)):
        plans, run, eager = sets[idx]
        out = g.run(plans, run["pts"], run["kdev"])
        got = {key: v.clone() for key, v in out.items() if v is not None}
        torch.cuda.synchronize()
        for key in eager:
            assert torch.equal(got[key], eager[key]), (rep, key)
        # eager work between the replays
        m._downstream(plans, run["pts"], run["kdev"])
        torch.cuda.synchronize()


def test_autotuner_picks_and_caches(tmp_path):
    """The autotuner times the candidates on a sample, applies the fastest,
    writes the on-disk cache and reuses it; results under the chosen shapes
    agree with the defaults (the MFMA lensing to rounding)."""
    import json
    from abcmb_amd import abi, autotune
    from abcmb_amd.model import Model, MassiveNeutrino
    m = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096)
    thetas = LK.latin_hypercube(4, seed=41)
    kg = [k[::4] for k in LK.uneven_k_grids(4, seed=41)]
    plans = m.prepare_many([m.derived(LK.theta_to_params(t)) for t in thetas], k_list=kg)
    ref = [o.ClTT.cpu().numpy() for o in m.execute(plans)]
    path = str(tmp_path / "tune.json")
    try:
        _say("case: autotuner on 4 cosmologies")
        choice = autotune.tune(m, plans, path=path)
        _say(f"autotuner choice: {choice}")
        assert choice["pert_persistent"] in autotune.PERT_CANDIDATES
        assert abi.hip_lib().abcmb_tune_get(abi.TUNE_PERT_PERSISTENT) == choice["pert_persistent"]
        assert m.lens_mode == choice["lens_mode"]
        data = json.load(open(path))
        assert len(data) == 1 and list(data.values())[0]["lens_mode"] == choice["lens_mode"]
        got = [o.ClTT.cpu().numpy() for o in m.execute(plans)]
        for a, b in zip(got, ref):
            assert np.max(np.abs(a / b - 1)) < 1e-11
        m2 = Model(user_species=(MassiveNeutrino,), lensing=True, l_max=2500, max_steps_PE=4096)
        again = autotune.tune(m2, plans, path=path)       # from the cache, no timing
        assert again == choice and m2.lens_mode == choice["lens_mode"]
    finally:
        abi.hip_lib().abcmb_tune_set(abi.TUNE_PERT_PERSISTENT, 0)
        abi.hip_lib().abcmb_tune_set(abi.TUNE_LENS_GEMM_COLS, 0)
