"""GPU parity on the benchmark workloads other than the batched one: the plain
forward dough block, the single differentiable rolling-pin scene, the
gripper + spatula gather scene with checkpointed adjoints and a Chamfer loss,
and the large knife-cutting scene (reduced against the oracle, full scale
through size-independent properties).  Each case runs through the drop-in
TaichiEnv / GradModel API and is compared with the C oracle on the same
seeded inputs: its fp32 build (the reference's ti.f32 arithmetic) holds the
bars, its fp64 build's distance is stated."""
import numpy as np
import pytest
import torch

from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg

pytestmark = pytest.mark.gpu


def gradmodel(w, ckpt_k=None, output_grid=True):
    from diffskill_amd.engine import GradModel, TaichiEnv
    cfg = dict(SIMULATOR=w["sim"], PRIMITIVES=w["prims"], SHAPES=[])
    env = TaichiEnv(cfg, init_particles=w["x0"], ckpt_k=ckpt_k)
    env.initialize()
    st = reset_state(w["x0"])
    n = st.shape[0]
    env.simulator.set_state(0, [st[:, 0:3], st[:, 3:6], st[:, 15:24].reshape(n, 3, 3), st[:, 6:15].reshape(n, 3, 3)]
                            + [prim_state_from_cfg(p) for p in w["prims"]])
    env.primitives.set_softness(666.0)
    return env, GradModel(env, softness=666.0, output_grid=(1,) if output_grid else ())


def flat_actions(w):
    a = w["actions"]
    return np.concatenate(a, axis=1) if isinstance(a, list) else np.asarray(a)


def oracle_rollout(w, H=None, bits=64, threads=16):
    osim, _ = sim_from_cfg(w["sim"])
    o = Oracle(osim, [prim_from_cfg(p) for p in w["prims"]], bits=bits, threads=threads)
    ro = OracleRollout(o, w["substeps"])
    acts = flat_actions(w)[:H]
    frames, ps = ro.run(reset_state(w["x0"]), np.stack([prim_state_from_cfg(p) for p in w["prims"]]), acts)
    return o, ro, frames, ps


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def target_grid(sim_cfg, x):
    """compute_grid_m of target particles through the engine."""
    from diffskill_amd.engine import MPMSimulator
    s = MPMSimulator(dict(sim_cfg, max_steps=2), ())
    s.reset(x)
    s.clear_and_compute_grid_m(0)
    return s.grid_m.to_torch().clone()


def test_config1_forward_parity():
    """5k-particle dough block, capsule, 50 env steps x 10 substeps at dt 2e-4
    (the 5e-4 of the workload statement diverges in the oracle as well):
    after all 500 substeps x within 1e-4 dx of the fp32 oracle (the
    reference's ti.f32 arithmetic), v within 1e-4 dx / dt (the velocity that
    moves a particle 1e-4 dx in one substep) and F within 1e-4; the distance
    of both to the fp64 oracle is printed."""
    from diffskill_amd import workloads
    w = workloads.config1(dt=2e-4)
    env, _ = gradmodel(w, output_grid=False)
    sim, P = env.simulator, env.primitives
    S, H = w["substeps"], w["H"]
    print("case config1 forward 500 substeps", flush=True)
    for s in range(H):
        P.set_action(s, S, w["actions"][s])
        for i in range(s * S, (s + 1) * S):
            sim.substep(i)
    o, ro, frames, ps = oracle_rollout(w)
    T = H * S
    f32 = oracle_rollout(w, bits=32)[2][T].astype(np.float64)
    x, v, F, C = sim.get_state(T)[:4]
    ref = frames[T]
    dx = 1.0 / 64
    F = F.reshape(-1, 9)
    print("vs fp32 oracle: x %.3g dx, v %.3g dx/dt, F %.3g; vs fp64: x %.3g dx (fp32 oracle %.3g dx)"
          % (np.abs(x - f32[:, 0:3]).max() / dx, np.abs(v - f32[:, 3:6]).max() * w["sim"]["dt"] / dx,
             np.abs(F - f32[:, 15:24]).max(), np.abs(x - ref[:, 0:3]).max() / dx, np.abs(f32[:, 0:3] - ref[:, 0:3]).max() / dx))
    assert np.isfinite(x).all()
    assert np.abs(x - f32[:, 0:3]).max() < 1e-4 * dx
    assert np.abs(v - f32[:, 3:6]).max() < 1e-4 * dx / w["sim"]["dt"]
    assert np.abs(F - f32[:, 15:24]).max() < 1e-4
    np.testing.assert_allclose(P[0].get_state(T), ps[T][0, :7], atol=1e-5)


def _grid_l1_grads(w, target, ckpt_k=None, H=None):
    env, gm = gradmodel(w, ckpt_k=ckpt_k)
    obs = gm.reset()
    acts = flat_actions(w)[:H]
    a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
    for s in range(len(acts)):
        obs = gm.forward(s, a[s], *obs)
    loss = (obs[2] - target).abs().sum()
    loss.backward()
    return float(loss.detach()), a.grad.detach().cpu().double().numpy()


def _oracle_grid_l1(w, target, H=None, bits=64):
    o, ro, frames, _ = oracle_rollout(w, H=H, bits=bits)
    T = len(frames) - 1
    m = o.grid_mass(frames[T]).astype(np.float64)
    tg = target.cpu().double().numpy()
    ga, _, _ = ro.grad({T: o.grid_mass_grad(frames[T], np.sign(m - tg).astype(o.dtype))})
    return float(np.abs(m - tg).sum()), np.asarray(ga, np.float64)


def test_config2_action_grads_full_horizon():
    """20k-particle ellipsoid, RollingPin, 500 substeps, L1 to the slab's mass
    grid: the action gradients through GradModel within 1e-3 (max element and
    L2, relative) of the fp32 oracle, the reference's ti.f32 arithmetic.
    Over this horizon fp32 rounding of the trajectory moves any fp32
    evaluation ~1.9e-3 from the fp64 oracle; that distance is printed."""
    from diffskill_amd import workloads
    w = workloads.config2()
    target = target_grid(w["sim"], w["target_x"])
    print("case config2 gpu", flush=True)
    loss, g = _grid_l1_grads(w, target)
    print("case config2 oracle", flush=True)
    loss_ref, g_ref = _oracle_grid_l1(w, target)
    loss32, g32 = _oracle_grid_l1(w, target, bits=32)
    l2 = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))  # noqa: E731
    print("loss %.6g vs %.6g (fp32 %.6g); grad vs fp32 oracle: max %.3g, l2 %.3g; vs fp64: max %.3g, l2 %.3g "
          "(fp32 oracle vs fp64: max %.3g, l2 %.3g)" % (loss, loss_ref, loss32, rel(g, g32), l2(g, g32), rel(g, g_ref),
                                                        l2(g, g_ref), rel(g32, g_ref), l2(g32, g_ref)))
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    assert rel(g, g32) < 1e-3
    assert l2(g, g32) < 1e-3


def _config3_grads(w, ckpt_k, H, det=True):
    from diffskill_amd.pointcloud import chamfer_distance
    from diffskill_amd.engine import GradModel, TaichiEnv
    cfg = dict(SIMULATOR=w["sim"], PRIMITIVES=w["prims"], SHAPES=[])
    env = TaichiEnv(cfg, init_particles=w["x0"], ckpt_k=ckpt_k, deterministic=det)
    env.initialize()
    st = reset_state(w["x0"])
    n = st.shape[0]
    env.simulator.set_state(0, [st[:, 0:3], st[:, 3:6], st[:, 15:24].reshape(n, 3, 3), st[:, 6:15].reshape(n, 3, 3)]
                            + [prim_state_from_cfg(p) for p in w["prims"]])
    env.primitives.set_softness(666.0)
    gm = GradModel(env, softness=666.0)
    obs = gm.reset()
    acts = flat_actions(w)[:H]
    a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
    for s in range(H):
        obs = gm.forward(s, a[s], *obs)
    t = torch.as_tensor(w["target_points"], dtype=torch.float32, device="cuda")
    # one-way Chamfer L2 (mean squared distance to the nearest target) on the GPU
    loss = chamfer_distance(obs[0][:, :3], t)
    loss.backward()
    ws = env.simulator.engine.ws.numel()
    return float(loss.detach()), a.grad.detach().cpu().double().numpy(), ws


def test_config3_checkpointed_equals_store_all():
    """Gather scene (30k particles, gripper + spatula, 1000 substeps, Chamfer
    loss) in the engine's fixed-order mode: checkpoints every K in {10, 50,
    100} frames recompute the store-all trajectory bitwise, so loss and
    action gradients equal the store-all run exactly (the reference's own
    check of its checkpointed recompute asks max |grad difference| < 1e-4,
    long_term_gradient.ipynb:255-257), in under half its device memory.  The
    default mode (unordered float atomics) is printed beside it: its rounding
    differs from run to run, which 1000 contact-rich substeps amplify to
    ~1e-3 relative in these gradients."""
    from diffskill_amd import workloads
    w = workloads.config3()
    H = w["H"]
    print("case config3 store-all (fixed order)", flush=True)
    loss0, g0, ws0 = _config3_grads(w, None, H)
    assert np.isfinite(g0).all() and np.abs(g0).max() > 0
    for K in (10, 50, 100):
        print("case config3 ckpt_k=%d (fixed order)" % K, flush=True)
        loss, g, ws = _config3_grads(w, K, H)
        print("K %d: workspace %.2f GB vs store-all %.2f GB, loss %.9g vs %.9g, max |grad difference| %.3g"
              % (K, ws / 1e9, ws0 / 1e9, loss, loss0, np.abs(g - g0).max()))
        assert loss == loss0
        assert np.array_equal(g, g0)
        assert ws < 0.5 * ws0
    lossf, gf, _ = _config3_grads(w, None, H, det=False)
    print("default mode vs fixed order: loss %.9g vs %.9g, grad rel %.3g" % (lossf, loss0, rel(gf, g0)))
    assert np.isclose(lossf, loss0, rtol=1e-5)
    assert rel(gf, g0) < 5e-3


def test_config3_oracle_parity_short_horizon():
    """Gather scene, first 2 env steps (40 substeps), all 30k particles: the
    Chamfer-loss action gradients against the fp64 oracle (nearest targets
    from a k-d tree on the CPU)."""
    from scipy.spatial import cKDTree
    from diffskill_amd import workloads
    w = workloads.config3()
    H = 2
    print("case config3 gpu short", flush=True)
    loss, g, _ = _config3_grads(w, None, H)
    o, ro, frames, _ = oracle_rollout(w, H=H)
    T = len(frames) - 1
    x = frames[T][:, 0:3]
    t = np.asarray(w["target_points"], np.float32).astype(np.float64)
    _, j = cKDTree(t).query(x)
    diff = x - t[j]
    loss_ref = float((diff ** 2).sum(axis=1).mean())
    adj = np.zeros_like(frames[T])
    adj[:, 0:3] = 2.0 * diff / x.shape[0]
    ga, _, _ = ro.grad({T: adj})
    print("loss %.8g vs %.8g, grad rel %.3g" % (loss, loss_ref, rel(g, ga)))
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert rel(g, ga) < 1e-3


def test_config5_reduced_parity():
    """Knife-cutting scene on the 128^3 grid, 27k particles (the block shrunk
    at the same particle density), 2 env steps of 50 substeps with checkpoints
    every 50 frames, mass-grid L1 loss: action gradients within 1e-3 of the
    fp32 oracle.  The across-blade gradients of this 2 mm plate are
    sensitive to fp32 rounding (the fp32 oracle is 3% from the fp64 one);
    that distance is printed."""
    from diffskill_amd import workloads
    w = workloads.config5(n=27000, n_env_steps=2)
    target = target_grid(w["sim"], w["target_x"])
    print("case config5 reduced gpu", flush=True)
    loss, g = _grid_l1_grads(w, target, ckpt_k=50)
    loss_ref, g_ref = _oracle_grid_l1(w, target)
    _, g32 = _oracle_grid_l1(w, target, bits=32)
    print("loss %.6g vs %.6g, grad rel vs fp32 oracle %.3g, vs fp64 %.3g (fp32 oracle vs fp64 %.3g)"
          % (loss, loss_ref, rel(g, g32), rel(g, g_ref), rel(g32, g_ref)))
    assert abs(loss - loss_ref) < 1e-4 * loss_ref
    assert rel(g, g32) < 1e-3


def test_config5_full_scale_properties():
    """The full cutting scene (1M particles, 128^3 grid, 2000 substeps, a
    checkpoint every 50 frames), forward + adjoint once: finite loss and
    gradients, total grid mass equal to N * p_mass at the last frame, the
    knife at its commanded depth, every particle inside the clamp box."""
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    w = workloads.config5()
    H, S = w["H"], w["substeps"]
    target = target_grid(w["sim"], w["target_x"])[None]
    b = SceneBatch([dict(sim=w["sim"], prims=w["prims"], x0=w["x0"])], H * S + 1, ckpt_k=w["ckpt_k"])
    init = torch.zeros((H, 1, 7))
    init[:, 0, :3] = torch.as_tensor(w["actions"], dtype=torch.float32)
    opt = BatchedTrajOpt(b, H, S, target, lr=0.0, init_actions=init, use_graph=False)
    print("case config5 full scale", flush=True)
    opt._device_pass()
    torch.cuda.synchronize()
    loss = float(opt.losses()[0])
    g = opt.action_grads()[:, 0, :3].cpu().double().numpy()
    assert np.isfinite(loss) and np.isfinite(g).all() and np.abs(g).max() > 0
    p_mass = (0.5 / 128) ** 2
    n = len(w["x0"])
    assert abs(float(opt.mass.sum()) - n * p_mass) < 1e-4 * n * p_mass
    knife = b.engine.prim_state[H * S, 0].cpu().double().numpy()
    y0 = w["prims"][0]["init_pos"][1]
    expect = max(y0 - 0.5 * 1e-4 * H * S, w["prims"][0]["lower_bound"][1])
    assert abs(knife[1] - expect) < 1e-4
    x = b.engine.frame(H * S)[0:3, b.engine.scene_cols(0)].t().cpu().numpy()
    assert np.isfinite(x).all() and x.min() >= 0.0 and x.max() <= 1.0 - 3.0 / 128 + 1e-6
    print("loss %.6g" % loss)
