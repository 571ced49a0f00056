"""GPU tests of the learned warm start: the BF16 MFMA network against its CPU reference
(oracle/mlp_ref.py, same rounding points), the feature kernel bit for bit, training,
the charges-x0 solve mode, and the gain of the trained network over a zero guess."""
import numpy as np
import pytest

import mlp_ref as R
from helpers import rel

pytestmark = pytest.mark.gpu

WORK = dict(xi=0.6, errortol=1e-6, rc=4.0, mesh=(27, 27, 27), P=8, restart=30, gmres_tol=1e-6,
            preconditioner="block", precision="fp32")


def _mlp(seed=5):
    from conductorrigid_amd.warmstart import WarmStartMLP
    return WarmStartMLP(device=0, seed=seed)


def _rows(M, seed):
    import torch
    rng = np.random.default_rng(seed)
    X = R.bf16(rng.uniform(-1, 1, size=(M, R.IN)).astype(np.float32))
    X[:, R.GATE + 6:] = 0.0
    X[:, R.GATE + 3:R.GATE + 6] *= 2.0 ** -9  # gate lo parts are small
    bits = R.bf16_bits(X).view(np.int16)
    return torch.from_numpy(bits.copy()).cuda(), X.astype(np.float64)


@pytest.mark.parametrize("M", [1, 16, 200, 1000])
def test_forward_matches_reference(require_gpu, M):
    m = _mlp()
    Xd, X = _rows(M, M)
    y = m.predict(Xd).cpu().numpy()
    yr, _ = R.forward(m.get_params(), X, bf16_points=True)
    assert rel(y, yr) < 2e-3, rel(y, yr)


def test_gradient_matches_reference(require_gpu):
    import torch
    m = _mlp(7)
    M = 333
    Xd, X = _rows(M, 11)
    t = np.random.default_rng(2).normal(scale=0.1, size=M).astype(np.float32)
    g, y = m.gradient(Xd, torch.from_numpy(t).cuda())
    g = g.cpu().numpy().astype(np.float64)
    gr, yr = R.gradient(m.get_params(), X, t.astype(np.float64), bf16_points=True)
    assert rel(y.cpu().numpy(), yr) < 2e-3
    for a, b in [(R.OFF_W1, R.OFF_W1 + R.TRUNK * R.H), (R.OFF_B1, R.OFF_W2), (R.OFF_W2, R.OFF_B2),
                 (R.OFF_B2, R.OFF_W3), (R.OFF_W3, R.OFF_B3), (R.OFF_B3, R.PARAMS)]:
        assert rel(g[a:b], gr[a:b]) < 5e-3, (a, rel(g[a:b], gr[a:b]))
    assert not np.any(g[R.OFF_W1 + R.TRUNK * R.H:R.OFF_B1])  # gate columns do not feed the trunk


def _cr(seed=21, nb=50, **kw):
    from conductorrigid_amd.compute import ConductorRigid
    from conductorrigid_amd.system import make_system
    s = make_system(nb, 20, 0.3, seed=seed, charges="uniform", n_isolated_beads=kw.pop("iso", 0))
    opts = dict(WORK)
    opts.update(kw)
    return ConductorRigid(system=s, field=(0.3, -0.2, 1.1), **opts)


def test_features_match_reference(require_gpu):
    cr = _cr(nb=20, iso=6)
    m = _mlp()
    X = m.features(cr).cpu().numpy().view(np.uint16)
    s = cr.system
    head, nn, nl = cr.dsys.nlist_host()
    Xr = R.features(s.pos, s.L, np.array([0.3, -0.2, 1.1]), s.iso_membership(), s.body_charge, s.rigid_membership,
                    s.rigid_com, head, nn, nl)
    Xr_bits = R.bf16_bits(Xr.astype(np.float32))
    assert np.array_equal(X, Xr_bits)
    assert np.count_nonzero(X[:, 3:160:5]) > 0  # some neighbours present
    g = R.gates(R.bf16_to_f32(X).astype(np.float64))
    assert np.allclose(g[:, :3], [0.3, -0.2, 1.1], rtol=1e-4, atol=0)  # hi + lo carries E0 to ~16 bits


def test_x0_charges_from_solution(require_gpu):
    cr = _cr(gmres_tol=1e-10, precision="fp64")
    st = cr.computeForces(0, x0="zero")
    q, Phi, F = cr.get_charges(), cr.get_potentials(), cr.get_forces()
    st2 = cr.computeForces(0, x0="charges")
    assert st2["iterations"] <= 1 and st2["converged"]
    assert rel(cr.get_charges(), q) < 1e-8 and rel(cr.get_potentials(), Phi) < 1e-8
    assert rel(cr.get_forces()[:, :3], F[:, :3]) < 1e-8
    # a perturbed q is projected onto the body charges before the solve
    import torch
    cr.charges.add_(torch.randn_like(cr.charges) * 0.01)
    st3 = cr.computeForces(0, x0="charges")
    assert st3["converged"] and rel(cr.get_charges(), q) < 1e-6


def test_training_reduces_loss(require_gpu):
    from conductorrigid_amd.warmstart import fit
    cr = _cr(seed=3)
    m = _mlp(1)
    pairs = []
    for E in (0.5, 1.0, 2.0):
        cr.update_field((0.0, 0.0, E))
        cr.computeForces(0, x0="zero")
        pairs.append((m.features(cr), m.targets(cr)))
    hist = fit(m, pairs, epochs=40, batch=512, lr=2e-3)
    print(f"\n[mlp] training loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.2 * hist[0]


def test_learned_warm_start_gain(require_gpu):
    """Train on solver outputs along a Brownian trajectory (random |E0| per step), then
    on held-out later steps compare GMRES iterations from four starts: zero, the previous
    step's solution (the reference's x0), the network's prediction, and the previous
    solution plus the network's predicted response to the field change."""
    import torch
    from conductorrigid_amd.warmstart import WarmStartMLP, bd_step, fit, run_trajectory
    cr = _cr(seed=8, nb=100, mesh=(32, 32, 32))
    m = WarmStartMLP(device=0, seed=3)
    n_train, n_eval = 90, 16
    recs, pairs = run_trajectory(cr, n_train, dt=2e-4, collect=range(n_train), mlp=m, seed=4)
    hist = fit(m, pairs, epochs=30, batch=2048, lr=2e-3)
    rng = np.random.default_rng(99)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(99)
    E_prev = 1.0
    cr.update_field((0.0, 0.0, E_prev))
    cr.computeForces(0, x0="zero")
    it = {"zero": [], "previous": [], "learned": [], "learned+previous": []}
    for k in range(n_eval):
        bd_step(cr, 2e-4, generator=gen)
        q_prev, P_prev = cr.charges.clone(), cr.potentials.clone()
        E = float(rng.uniform(0.5, 2.0))
        cr.update_field((0.0, 0.0, E))
        it["zero"].append(cr.computeForces(k, x0="zero")["iterations"])
        q_ref = cr.get_charges()
        cr.charges.copy_(q_prev)
        cr.potentials.copy_(P_prev)
        it["previous"].append(cr.computeForces(k, x0="previous")["iterations"])
        m.warm_start(cr)
        it["learned"].append(cr.computeForces(k, x0="charges")["iterations"])
        assert rel(cr.get_charges(), q_ref) < 1e-4  # same answer to the solve tolerance
        cr.charges.copy_(q_prev)
        m.warm_start_from_previous(cr, (0.0, 0.0, E_prev))
        st = cr.computeForces(k, x0="charges")
        it["learned+previous"].append(st["iterations"])
        assert st["converged"] and rel(cr.get_charges(), q_ref) < 1e-4
        E_prev = E
    mean = {k: float(np.mean(v)) for k, v in it.items()}
    print(f"\n[warm start] training loss {hist[0]:.3e} -> {hist[-1]:.3e}; mean GMRES iterations " +
          ", ".join(f"{k} {v:.1f}" for k, v in mean.items()))
    # the network's start saves iterations over a zero guess; added to the previous
    # solution its field-response correction is within noise of the previous start
    assert mean["learned"] < mean["zero"]
    assert mean["learned+previous"] <= mean["previous"] + 1.0


def test_config5_workload_learned_start(require_gpu):
    """The learned-warm-start workload at its own size (100k beads, 5000 bodies x 20, the
    n100k parameters), on a short version of its protocol: a Brownian trajectory with
    |E0| ~ U[0.5, 2] per step, the network trained on solver outputs of the first steps,
    then held-out steps solved from zero and from the network's start.  The learned start
    gives the zero start's answer (to the solver tolerance) with no fallback."""
    import torch
    from conductorrigid_amd.compute import ConductorRigid
    from conductorrigid_amd.system import make_workload
    from conductorrigid_amd.warmstart import WarmStartMLP, bd_step, fit, run_trajectory

    s, w = make_workload("n100k")
    assert s.N == 100000 and s.N_rigid == 5000
    cr = ConductorRigid(system=s, field=(0.0, 0.0, 1.0), xi=w["xi"], errortol=w["errortol"], rc=w["rc"],
                        mesh=w["mesh"], P=w["P"], restart=w["restart"], gmres_tol=w["tol"], max_iter=2000,
                        precision="fp32", preconditioner="block")
    m = WarmStartMLP(device=0, seed=5)
    n_train, n_eval = 40, 4
    recs, pairs = run_trajectory(cr, n_train, dt=2e-4, collect=range(0, n_train, 4), mlp=m, seed=6)
    hist = fit(m, pairs, epochs=3, batch=8192, lr=2e-3)
    del pairs
    rng = np.random.default_rng(7)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    it = {"zero": [], "learned": []}
    for k in range(n_eval):
        bd_step(cr, 2e-4, generator=gen)
        cr.update_field((0.0, 0.0, float(rng.uniform(0.5, 2.0))))
        st0 = cr.computeForces(k, x0="zero")
        q_ref = cr.get_charges()
        m.warm_start(cr)
        st = cr.computeForces(k, x0="charges")
        assert st0["converged"] and st["converged"] and st["fallback"] == 0
        assert rel(cr.get_charges(), q_ref) < 1e-4
        it["zero"].append(st0["iterations"])
        it["learned"].append(st["iterations"])
    print(f"\n[config5] 100k beads: loss {hist[0]:.3e} -> {hist[-1]:.3e}; iterations zero {it['zero']}, "
          f"learned {it['learned']}", flush=True)
    assert hist[-1] < hist[0]
