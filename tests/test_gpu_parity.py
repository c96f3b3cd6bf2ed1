"""GPU parity: the HIP engine (through the C ABI and the drop-in Python API)
against the fp64 CPU oracle on the same seeded inputs."""
import numpy as np
import pytest
import torch

from oracle import Oracle, OracleRollout
from scenes import capsule_scene, oracle_prim, oracle_sim, particles_box, prim_state0, state_list

pytestmark = pytest.mark.gpu


def build(sim_cfg, prim_cfgs, st, softness=666.0, ckpt_k=None):
    from diffskill_amd.engine import MPMSimulator, Primitives
    prims = Primitives(prim_cfgs)
    sim = MPMSimulator(sim_cfg, prims, ckpt_k=ckpt_k)
    sim.set_state(0, state_list(st, [prim_state0(p) for p in prim_cfgs]))
    prims.set_softness(softness)
    return sim, prims


def flat(state):
    n = state[0].shape[0]
    out = np.zeros((n, 24))
    out[:, 0:3] = state[0]
    out[:, 3:6] = state[1]
    out[:, 6:15] = state[3].reshape(n, 9)
    out[:, 15:24] = state[2].reshape(n, 9)
    return out


def test_state_roundtrip_in_caller_order():
    """Particles are stored in cell order internally; get_state returns them in
    the caller's order, bit-exact after the fp32 round trip."""
    sim_cfg, prims, st = capsule_scene(n=3000, seed=8)
    st = particles_box(3000, [0.1, 0.1, 0.1], [0.9, 0.9, 0.9], seed=8, v_sigma=1.0, F_jitter=0.1)
    sim, P = build(sim_cfg, prims, st)
    got = flat(sim.get_state(0))
    np.testing.assert_array_equal(got, st.astype(np.float32).astype(np.float64))
    perm = sim._perm.cpu().numpy()
    assert not np.array_equal(perm, np.arange(3000))
    xs = sim.engine.frame(0)[0:3, sim.engine.scene_cols(0)].t().cpu().numpy()
    np.testing.assert_array_equal(xs, st[perm, 0:3].astype(np.float32))


def test_grid_mass_parity():
    sim_cfg, prims, st = capsule_scene(n=3000, seed=3)
    sim, P = build(sim_cfg, prims, st)
    print("case grid_mass n=3000", flush=True)
    sim.clear_and_compute_grid_m(0)
    m = sim.get_m()
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    m_ref = o.grid_mass(st)
    assert abs(m.sum() - 3000 * c["p_mass"]) < 1e-5 * 3000 * c["p_mass"]
    np.testing.assert_allclose(m, m_ref, rtol=1e-5, atol=1e-6 * m_ref.max())


@pytest.mark.parametrize("F_jitter", [0.0, 0.05])
def test_forward_substeps_parity(F_jitter):
    sim_cfg, prims, st = capsule_scene(n=2000, seed=1)
    st = particles_box(2000, [0.4, 0.05, 0.4], [0.6, 0.15, 0.6], seed=1, v_sigma=0.1, F_jitter=F_jitter)
    sim, P = build(sim_cfg, prims, st)
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    H, S = 4, c["substeps"]
    rng = np.random.default_rng(7)
    acts = rng.uniform(-0.8, 0.8, (H, 3))
    print("case forward F_jitter=%g" % F_jitter, flush=True)
    for s in range(H):
        P.set_action(s, S, acts[s])
        for i in range(s * S, (s + 1) * S):
            sim.substep(i)
    ro = OracleRollout(o, S)
    frames, ps = ro.run(st, prim_state0(prims[0])[None], acts)
    T = H * S
    got = flat(sim.get_state(T))
    ref = frames[T]
    dx = c["dx"]
    assert np.abs(got[:, 0:3] - ref[:, 0:3]).max() < 1e-4 * dx
    vs = np.abs(ref[:, 3:6]).max()
    assert np.abs(got[:, 3:6] - ref[:, 3:6]).max() < 1e-3 * vs
    assert np.abs(got[:, 15:24] - ref[:, 15:24]).max() < 1e-4
    np.testing.assert_allclose(sim.primitives[0].get_state(T), ps[T][0, :7], atol=1e-6)


def test_adjoint_single_substep_parity():
    sim_cfg, prims, st = capsule_scene(n=1500, seed=2)
    st = particles_box(1500, [0.4, 0.05, 0.4], [0.6, 0.15, 0.6], seed=2, v_sigma=0.1, F_jitter=0.08)
    sim, P = build(sim_cfg, prims, st)
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    S = c["substeps"]
    a0 = np.array([0.3, -0.5, 0.2])
    P.set_action(0, S, a0)
    sim.substep(0)
    rng = np.random.default_rng(11)
    adj1 = rng.normal(size=(1500, 24))
    sim.clear_grads()
    import torch as T_
    xv = T_.as_tensor(adj1[:, 0:6], dtype=T_.float32)
    sim.add_obs_grad(1, xv, None)
    # the C and F parts of the injected adjoint go straight into the engine's adjoint frame
    a = sim.engine.adjoint(1)
    cols = sim.engine.scene_cols(0)
    cf = sim._to_slots(T_.as_tensor(adj1[:, 6:24], dtype=T_.float32, device="cuda"))
    a[6:24, cols] = cf.t()
    sim.substep_grad(0)
    gx, gv, gF, gC = [t.detach().cpu().double().numpy() for t in sim.get_adjoint(0)]
    pv = np.zeros((1, 7))
    pv[0] = o.set_velocity(0, a0, S)
    ps0 = prim_state0(prims[0])[None]
    st1, ps1 = o.substep(st, ps0, pv)
    ref0, gps0, gps1, gpv = o.substep_grad(st, adj1, ps0, pv, ps1, np.zeros((1, 8)))
    got = np.concatenate([gx, gv, gC.reshape(-1, 9), gF.reshape(-1, 9)], 1)
    for name, sl in (("x", slice(0, 3)), ("v", slice(3, 6)), ("C", slice(6, 15)), ("F", slice(15, 24))):
        err = np.linalg.norm(got[:, sl] - ref0[:, sl]) / np.linalg.norm(ref0[:, sl])
        print(name, err)
        assert err < 1e-3, name
    gps_gpu = sim.engine.prim_state_grad[0, 0].cpu().double().numpy()
    gpv_gpu = sim.engine.prim_vel_grad[0, 0].cpu().double().numpy()
    np.testing.assert_allclose(gps_gpu[:7], gps0[0, :7], rtol=2e-3, atol=2e-3 * np.abs(gps0).max())
    np.testing.assert_allclose(gpv_gpu[:6], gpv[0, :6], rtol=2e-3, atol=2e-3 * np.abs(gpv).max())


def _gradmodel_loss_grads(sim_cfg, prims, st, acts, target, ckpt_k=None):
    from diffskill_amd.engine import GradModel, TaichiEnv
    cfg = dict(SIMULATOR=sim_cfg, PRIMITIVES=prims, SHAPES=[])
    env = TaichiEnv(cfg, init_particles=st[:, :3], ckpt_k=ckpt_k)
    env.initialize()
    env.simulator.set_state(0, state_list(st, [prim_state0(p) for p in prims]))
    env.primitives.set_softness(666.0)
    gm = GradModel(env, softness=666.0, output_grid=(1,))
    obs = gm.reset()
    a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
    tgt = torch.as_tensor(target, dtype=torch.float32, device="cuda")
    for s in range(len(acts)):
        obs = gm.forward(s, a[s], *obs)
    loss = (obs[2] - tgt).abs().sum() + 1e-3 * (obs[0] ** 2).sum()
    loss.backward()
    return float(loss.detach()), a.grad.detach().cpu().double().numpy()


def _oracle_loss_grads(sim_cfg, prims, st, acts, target):
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    ro = OracleRollout(o, c["substeps"])
    frames, _ = ro.run(st, np.stack([prim_state0(p) for p in prims]), acts)
    T = len(frames) - 1
    m = o.grid_mass(frames[T])
    loss = np.abs(m - target).sum() + 1e-3 * (frames[T][:, 0:6] ** 2).sum()
    adj = o.grid_mass_grad(frames[T], np.sign(m - target))
    adj[:, 0:6] += 2e-3 * frames[T][:, 0:6]
    ga, _, _ = ro.grad({T: adj})
    return loss, ga


@pytest.mark.parametrize("ckpt_k", [None, 4])
def test_gradmodel_action_grad_parity(ckpt_k):
    sim_cfg, prims, st = capsule_scene(n=2000, seed=4)
    rng = np.random.default_rng(5)
    acts = rng.uniform(-0.6, 0.6, (3, 3))
    n = sim_cfg["n_grid"]
    target = np.random.default_rng(6).uniform(0, 2e-4, (n, n, n))
    print("case gradmodel ckpt_k=%s" % ckpt_k, flush=True)
    loss, g = _gradmodel_loss_grads(sim_cfg, prims, st, acts, target, ckpt_k=ckpt_k)
    loss_ref, g_ref = _oracle_loss_grads(sim_cfg, prims, st, acts, target)
    print(loss, loss_ref)
    print(g)
    print(g_ref)
    assert abs(loss - loss_ref) < 1e-4 * abs(loss_ref)
    rel = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    assert rel < 1e-3, rel


def test_field_views_match_state_and_adjoint_accessors():
    """The Taichi-field views callers reach into (mpm_simulator.py:40-43,
    primive_base.py:37-52, function.py:44-61): particle x/v/F/C[f, p],
    primitive position/rotation/v/w/gap/action_buffer[f] and their .grad,
    against get_state, the action table and GradModel's action gradients."""
    from diffskill_amd.engine import GradModel, TaichiEnv
    print("case field views", flush=True)
    sim_cfg, prims, st = capsule_scene(n=1000, seed=8)
    prims = prims + [dict(shape="Gripper", size=(0.01, 0.03, 0.03), minimal_gap=0.02, init_gap=0.12, init_pos=(0.3, 0.3, 0.3),
                          init_rot=(1.0, 0.0, 0.0, 0.0), action=dict(dim=7, scale=(0.01,) * 7))]
    env = TaichiEnv(dict(SIMULATOR=sim_cfg, PRIMITIVES=prims, SHAPES=[]), init_particles=st[:, :3])
    env.initialize()
    env.simulator.set_state(0, state_list(st, [prim_state0(p) for p in prims]))
    env.primitives.set_softness(666.0)
    gm = GradModel(env, softness=666.0, output_grid=(1,))
    obs = gm.reset()
    sim, P = env.simulator, env.primitives
    A = P.action_dims[-1]
    a = torch.tensor(np.random.default_rng(1).uniform(-0.5, 0.5, (2, A)), dtype=torch.float32, device="cuda", requires_grad=True)
    for s in range(2):
        obs = gm.forward(s, a[s], *obs)
    S = sim.substeps
    f = 2 * S
    state = sim.get_state(f)
    for p in (0, 17, 999):
        np.testing.assert_allclose(sim.x[f, p], state[0][p], rtol=0, atol=0)
        np.testing.assert_allclose(sim.v[f, p], state[1][p], rtol=0, atol=0)
        np.testing.assert_allclose(sim.F[f, p], state[2][p], rtol=0, atol=0)
        np.testing.assert_allclose(sim.C[f, p], state[3][p], rtol=0, atol=0)
    xs = sim.x.to_numpy()
    assert xs.shape == (sim.max_steps, 1000, 3)
    np.testing.assert_allclose(xs[f], state[0], rtol=0, atol=0)
    g = P[1]
    np.testing.assert_allclose(np.r_[g.position[f], g.rotation[f], g.gap[f]], g.get_state(f), rtol=0, atol=0)
    np.testing.assert_allclose(g.action_buffer[1], np.clip(a[1, 3:].detach().cpu().numpy(), -1, 1), atol=1e-7)
    # v, w, gap_vel of a substep: the action times its scale over the substeps (set_velocity)
    act = np.clip(a[0, 3:].detach().cpu().numpy(), -1, 1)
    np.testing.assert_allclose(g.v[1], act[:3] * 0.01 / S, atol=1e-8)
    np.testing.assert_allclose(g.w[1], act[3:6] * 0.01 / S, atol=1e-8)
    assert abs(g.gap_vel[1] - act[6] * 0.01 / S) < 1e-8
    loss = (obs[2] ** 2).sum() * 1e3 + (obs[1] ** 2).sum()
    loss.backward()
    ga = a.grad.detach().cpu().double().numpy()
    np.testing.assert_allclose(g.action_buffer.grad[0], ga[0, 3:], rtol=1e-6, atol=1e-12)
    assert np.abs(g.position.grad.to_numpy()[0]).max() > 0
    g.position[0] = [0.31, 0.32, 0.33]
    np.testing.assert_allclose(g.get_state(0)[:3], [0.31, 0.32, 0.33], atol=1e-7)
    sim.x.grad.fill(0)
    assert np.abs(g.position.grad.to_numpy()).max() == 0


@pytest.mark.parametrize("corner", ["low", "high"])
def test_walls_and_clamp_parity(corner):
    """Particles driven into the domain walls (grid_op's bound-3 wall
    projection and ground Coulomb friction, mpm_simulator.py:241-260; g2p's
    position clamp, :283): 20 forward substeps against the fp64 oracle, and
    one adjoint substep taken at the walls."""
    sim_cfg, prims, _ = capsule_scene(n=1500, seed=12)
    if corner == "low":
        st = particles_box(1500, [0.08, 0.08, 0.08], [0.14, 0.14, 0.14], seed=12, v_sigma=0.5, F_jitter=0.03)
        st[:, 3:6] += -8.0
    else:
        st = particles_box(1500, [0.85, 0.85, 0.85], [0.9, 0.9, 0.9], seed=12, v_sigma=0.5, F_jitter=0.03)
        st[:, 3:6] += 8.0
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    dx, S, H = c["dx"], c["substeps"], 4
    acts = np.zeros((H, 3))
    print("case walls %s forward" % corner, flush=True)
    sim, P = build(sim_cfg, prims, st)
    for s in range(H):
        P.set_action(s, S, acts[s])
        for i in range(s * S, (s + 1) * S):
            sim.substep(i)
    frames, _ = OracleRollout(o, S).run(st, prim_state0(prims[0])[None], acts)
    T = H * S
    got, ref = flat(sim.get_state(T)), frames[T]
    # the walls were reached: some particle's stencil touches a wall node
    lo_n = np.floor(ref[:, 0:3] / dx - 0.5)
    assert (lo_n.min() < 3) if corner == "low" else (lo_n.max() + 2 > c["n_grid"] - 3)
    if corner == "high":   # and positions were clamped at 1 - 3 dx
        assert any(np.any(fr[:, 0:3] >= 1 - 3 * dx - 1e-12) for fr in frames)
    assert np.abs(got[:, 0:3] - ref[:, 0:3]).max() < 1e-4 * dx
    assert np.abs(got[:, 3:6] - ref[:, 3:6]).max() < 1e-3 * np.abs(ref[:, 3:6]).max()
    assert np.abs(got[:, 15:24] - ref[:, 15:24]).max() < 1e-4

    print("case walls %s adjoint" % corner, flush=True)
    sw = frames[T]
    sim, P = build(sim_cfg, prims, sw)
    P.set_action(0, S, acts[0])
    sim.substep(0)
    adj1 = np.random.default_rng(13).normal(size=(1500, 24))
    sim.clear_grads()
    sim.add_obs_grad(1, torch.as_tensor(adj1[:, 0:6], dtype=torch.float32), None)
    a = sim.engine.adjoint(1)
    a[6:24, sim.engine.scene_cols(0)] = sim._to_slots(torch.as_tensor(adj1[:, 6:24], dtype=torch.float32, device="cuda")).t()
    sim.substep_grad(0)
    gx, gv, gF, gC = [t.detach().cpu().double().numpy() for t in sim.get_adjoint(0)]
    ps0 = prim_state0(prims[0])[None]
    pv = o.set_velocity(0, acts[0], S)[None]
    _, ps1 = o.substep(sw, ps0, pv)
    ref0, _, _, _ = o.substep_grad(sw, adj1, ps0, pv, ps1, np.zeros((1, 8)))
    got = np.concatenate([gx, gv, gC.reshape(-1, 9), gF.reshape(-1, 9)], 1)
    for name, sl in (("x", slice(0, 3)), ("v", slice(3, 6)), ("C", slice(6, 15)), ("F", slice(15, 24))):
        err = np.linalg.norm(got[:, sl] - ref0[:, sl]) / np.linalg.norm(ref0[:, sl])
        print(name, err, flush=True)
        assert err < 1e-3, name


def test_copy_mode_step_matches_oracle_rollout():
    """MPMSimulator.step in copy mode (mpm_simulator.py:440-451): each env step
    runs its substeps from frame 0 and copies the last frame (particles and
    tools) back to frame 0, so after k steps frame 0 holds the state of frame
    k * substeps of one continuous oracle rollout with the same actions;
    non-copy steps advance cur and leave the frames in place."""
    sim_cfg, prims, st = capsule_scene(n=1500, seed=8, substeps=4)
    acts = np.random.default_rng(9).uniform(-0.8, 0.8, (3, 3))
    osim, c = oracle_sim(sim_cfg)
    o = Oracle(osim, [oracle_prim(p) for p in prims])
    frames, ps = OracleRollout(o, c["substeps"]).run(st, np.stack([prim_state0(p) for p in prims]), acts)
    S = c["substeps"]
    dx = 1.0 / sim_cfg["n_grid"]
    for is_copy in (True, False):
        print("case copy-mode step is_copy=%s" % is_copy, flush=True)
        sim, P = build(sim_cfg, prims, st)
        for k in range(len(acts)):
            sim.step(is_copy, acts[k])
            f = 0 if is_copy else (k + 1) * S
            assert sim.cur == f
            got = flat(sim.get_state(f))
            ref = frames[(k + 1) * S]
            assert np.abs(got[:, 0:3] - ref[:, 0:3]).max() < 1e-4 * dx, (is_copy, k)
            assert np.abs(got[:, 15:24] - ref[:, 15:24]).max() < 1e-4, (is_copy, k)
            np.testing.assert_allclose(P[0].get_state(f), ps[(k + 1) * S][0, :7], atol=1e-5)
