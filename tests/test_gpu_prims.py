"""GPU parity per tool primitive class: the same short differentiable rollout
(2 env steps x 3 substeps, 32^3 grid, a tilted tool pressing into a dough
block) through GradModel on the HIP engine and through the fp64 oracle, for
every primitive shape the engine implements (SDF, normal -- analytic or
central-difference --, collider velocity, kinematics, gap DoF and all their
adjoints).  Each case checks the forward state and the action gradients."""
import numpy as np
import pytest

from test_gpu_parity import _gradmodel_loss_grads, _oracle_loss_grads, build, flat
from scenes import oracle_prim, oracle_sim, particles_box, prim_state0

from oracle import Oracle, OracleRollout

pytestmark = pytest.mark.gpu

SHAPES = {
    "Capsule": dict(shape="Capsule", h=0.1, r=0.03, action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Sphere": dict(shape="Sphere", radius=0.04, action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Box": dict(shape="Box", size=(0.03, 0.02, 0.04), action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
    "Prism": dict(shape="Prism", h=(0.05, 0.05), prot=(0.9, 0.1, 0.3, 0.2), action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Knife": dict(shape="Knife", h=(0.05, 0.05), size=(0.03, 0.02, 0.04), prot=(0.9, 0.1, 0.3, 0.2),
                  action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "RollingPin": dict(shape="RollingPin", h=0.1, r=0.03, action=dict(dim=3, scale=(0.5, 0.05, 0.002))),
    "RollingPinExt": dict(shape="RollingPinExt", h=0.1, r=0.03, action=dict(dim=6, scale=(0.5, 0.05, 0.002, 0.01, 0.01, 0.01))),
    "Gripper": dict(shape="Gripper", size=(0.01, 0.03, 0.03), minimal_gap=0.02, init_gap=0.12,
                    action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
    "Gripper2": dict(shape="Gripper2", h=0.06, r=0.015, minimal_gap=0.02, init_gap=0.12,
                     action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
    "Chopsticks": dict(shape="Chopsticks", h=0.1, r=0.02, minimal_gap=0.02, init_gap=0.1,
                       action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
}


def scene(shape):
    sim = dict(yield_stress=50.0, E=5e3, nu=0.2, ground_friction=1.5, gravity=(0, -9.8, 0), n_grid=32, dt=2e-4,
               substeps=3, max_steps=16, n_particles=2000)
    cfg = dict(SHAPES[shape])
    q = np.array([0.9, 0.2, 0.3, 0.1])
    q /= np.linalg.norm(q)
    cfg.update(init_pos=(0.5, 0.185, 0.5), init_rot=tuple(q), friction=0.9)
    st = particles_box(2000, [0.4, 0.05, 0.4], [0.6, 0.15, 0.6], seed=5, v_sigma=0.1)
    return sim, [cfg], st


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_primitive_forward_and_action_grads(shape):
    print("case primitive %s" % shape, flush=True)
    sim_cfg, prims, st = scene(shape)
    A = prims[0]["action"]["dim"]
    acts = np.random.default_rng(3).uniform(-0.5, 0.5, (2, A))
    # forward state after both env steps
    sim, P = build(sim_cfg, prims, st)
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    S = c["substeps"]
    for s in range(2):
        P.set_action(s, S, acts[s])
        for i in range(s * S, (s + 1) * S):
            sim.substep(i)
    frames, ps = OracleRollout(o, S).run(st, prim_state0(prims[0])[None], acts)
    got = flat(sim.get_state(2 * S))
    ref = frames[2 * S]
    assert np.abs(got[:, 0:3] - ref[:, 0:3]).max() < 1e-4 * c["dx"]
    assert np.abs(got[:, 3:6] - ref[:, 3:6]).max() < 1e-3 * np.abs(ref[:, 3:6]).max()
    np.testing.assert_allclose(P[0].get_state(2 * S), ps[2 * S][0, :len(P[0].get_state(2 * S))], atol=1e-6)
    # gradients of a mass-grid L1 + L2 loss w.r.t. the actions
    target = np.random.default_rng(2).uniform(0, 1e-4, (32, 32, 32))
    loss, g = _gradmodel_loss_grads(sim_cfg, prims, st, acts, target)
    loss_ref, g_ref = _oracle_loss_grads(sim_cfg, prims, st, acts, target)
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print("%s: loss %.8g vs %.8g, grad rel %.3g" % (shape, loss, loss_ref, err))
    assert abs(loss - loss_ref) < 1e-5 * abs(loss_ref)
    assert err < 1e-3, (g, g_ref)
