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
    "Cylinder": dict(shape="Cylinder", h=0.04, r=0.02, action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
    "Torus": dict(shape="Torus", tx=0.05, ty=0.015, action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
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


@pytest.mark.parametrize("shape", sorted(s for s in SHAPES if s != "Chopsticks"))
def test_return_dist_observations_and_gradients(shape):
    """GradModel(return_dist=True): the per-particle tool distances appended
    to the observation (function.py:79-88,104-116) and their adjoint
    (function.py:154-160), injected at every observed step, against the fp64
    oracle."""
    import torch
    from diffskill_amd.engine import GradModel, TaichiEnv
    from scenes import state_list
    print("case return_dist %s" % shape, flush=True)
    sim_cfg, prims, st = scene(shape)
    A = prims[0]["action"]["dim"]
    acts = np.random.default_rng(3).uniform(-0.5, 0.5, (2, A))
    env = TaichiEnv(dict(SIMULATOR=sim_cfg, PRIMITIVES=prims, SHAPES=[]), init_particles=st[:, :3])
    env.initialize()
    env.simulator.set_state(0, state_list(st, [prim_state0(p) for p in prims]))
    env.primitives.set_softness(666.0)
    gm = GradModel(env, softness=666.0, output_grid=(1,), return_dist=True)
    obs = gm.reset()
    nd = obs[0].shape[1] - 6
    assert nd == (2 if shape in ("Gripper", "Gripper2") else 1)
    w = [torch.as_tensor(np.random.default_rng(10 + s).normal(size=(2000, nd)), dtype=torch.float32, device="cuda")
         for s in range(2)]
    a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
    loss = 0
    d_gpu = []
    for s in range(2):
        obs = gm.forward(s, a[s], *obs)
        d_gpu.append(obs[0][:, 6:].detach().cpu().double().numpy())
        loss = loss + (w[s] * obs[0][:, 6:]).sum()
    loss = loss + 1e-3 * obs[2].sum()
    loss.backward()
    g = a.grad.detach().cpu().double().numpy()

    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    S = c["substeps"]
    ro = OracleRollout(o, S)
    frames, ps = ro.run(st, prim_state0(prims[0])[None], acts)
    frame_adj, prim_adj = {}, {}
    for s in range(2):
        f = (s + 1) * S
        d_ref = o.min_dist(frames[f], ps[f]).T
        np.testing.assert_allclose(d_gpu[s], d_ref, rtol=0, atol=1e-6)
        adj, gps = o.min_dist_grad(frames[f], ps[f], w[s].cpu().double().numpy().T)
        frame_adj[f] = adj
        prim_adj[f] = gps
    T = 2 * S
    frame_adj[T] = frame_adj[T] + o.grid_mass_grad(frames[T], np.full((32, 32, 32), 1e-3))
    ga, _, _ = ro.grad(frame_adj, prim_adj)
    err = np.abs(g - ga).max() / np.abs(ga).max()
    print("%s: return_dist grad rel %.3g" % (shape, err))
    assert err < 1e-3, (g, ga)


def test_taichi_env_return_dist_columns():
    """TaichiEnv(return_dist=True).get_obs: positions plus one distance column
    per primitive (taichi_env.py:85-98,356-383): both jaws summed for a
    Gripper, Gripper2.sdf (nearer jaw) and Sphere.sdf otherwise."""
    from diffskill_amd.engine import TaichiEnv
    from scenes import state_list
    sim_cfg, _, st = scene("Sphere")
    prims = []
    for name, pos in (("Gripper", (0.5, 0.2, 0.5)), ("Gripper2", (0.45, 0.16, 0.55)), ("Sphere", (0.55, 0.19, 0.45))):
        cfg = dict(SHAPES[name])
        cfg.update(init_pos=pos, init_rot=(1.0, 0.0, 0.0, 0.0), friction=0.9)
        prims.append(cfg)
    env = TaichiEnv(dict(SIMULATOR=sim_cfg, PRIMITIVES=prims, SHAPES=[]), init_particles=st[:, :3], return_dist=True)
    env.initialize()
    ps0 = [prim_state0(p) for p in prims]
    env.simulator.set_state(0, state_list(st, ps0))
    x, c = env.get_obs(0, "cuda")
    assert x.shape == (2000, 6)
    osim, _ = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    d = o.min_dist(st, np.stack(ps0))          # Gripper jaws, Gripper2 jaws, Sphere
    ref = np.stack([d[0] + d[1], np.minimum(d[2], d[3]), d[4]], 1)
    np.testing.assert_allclose(x[:, 3:].cpu().double().numpy(), ref, rtol=0, atol=1e-6)
    np.testing.assert_allclose(x[:, :3].cpu().double().numpy(), st[:, :3], rtol=0, atol=1e-7)
