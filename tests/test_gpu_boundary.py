"""GPU checks of the drop-in boundary's limits and error paths:

- a scene with six tools, the task file plb/envs/triplemove.yml (three dough
  boxes of 3333 particles, six Spheres; its values restated here), loads
  through TaichiEnv and a short GradModel rollout matches the oracle; so does
  a scene with eight Spheres, the most collision units a scene carries;
- a scene with more collision units than the engine carries is refused with
  a clear error;
- the entry points that add into a frame's adjoint refuse frames outside the
  context or not yet simulated, and such a refusal leaves the adjoint chain
  as it was."""
import numpy as np
import pytest
import torch

from oracle import Oracle, OracleRollout
from scenes import oracle_prim, oracle_sim, prim_state0, state_list

pytestmark = pytest.mark.gpu


def triplemove_cfg():
    """plb/envs/triplemove.yml:1-64 (SIMULATOR, SHAPES, PRIMITIVES)."""
    shapes = [dict(shape="box", width=(0.1, 0.1, 0.1), init_pos=(x, 0.05, 0.5), n_particles=3333) for x in (0.3, 0.5, 0.7)]
    prims = [dict(shape="Sphere", radius=0.03, init_pos=(x, 0.05, 0.5), color=(0.8, 0.8, 0.8), friction=0.9,
                  action=dict(dim=3, scale=(0.01, 0.01, 0.01))) for x in (0.23, 0.37, 0.43, 0.57, 0.63, 0.77)]
    return dict(SIMULATOR=dict(yield_stress=200.0), SHAPES=shapes, PRIMITIVES=prims)


def eight_sphere_cfg():
    """Eight Spheres (eight collision units, the most a scene carries: the
    grid kernels stage that many primitives per scene), four around each of
    two dough boxes."""
    shapes = [dict(shape="box", width=(0.1, 0.1, 0.1), init_pos=(x, 0.05, 0.5), n_particles=1500) for x in (0.35, 0.65)]
    pos = []
    for x in (0.35, 0.65):
        pos += [(x - 0.075, 0.05, 0.5), (x + 0.075, 0.05, 0.5), (x, 0.05, 0.425), (x, 0.05, 0.575)]
    prims = [dict(shape="Sphere", radius=0.03, init_pos=p, color=(0.8, 0.8, 0.8), friction=0.9,
                  action=dict(dim=3, scale=(0.01, 0.01, 0.01))) for p in pos]
    return dict(SIMULATOR=dict(yield_stress=200.0), SHAPES=shapes, PRIMITIVES=prims)


def rollout_matches_oracle(cfg, n_particles, n_tools, seed):
    """A short GradModel rollout of the task (2 env steps, density-grid L1
    loss) against the fp32 / fp64 oracle: final positions, loss and action
    gradients; every tool's actions carry gradient.  Positions and gradients
    are held to the fp32 oracle (the reference's ti.f32 arithmetic), with the
    fp64 distances reported: over 38 substeps of eight spheres pressing the
    dough, fp32 rounding alone moves particles by up to 1.2e-4 dx from the
    fp64 trajectory (the fp32 oracle's own distance), past the 1e-4 dx bound.
    The dough clouds come from numpy's global generator, seeded here so that
    each run samples the same particles."""
    from diffskill_amd.engine import GradModel, TaichiEnv
    np.random.seed(seed)
    env = TaichiEnv(cfg)
    env.initialize()
    sim = env.simulator
    n = env.n_particles
    assert n == n_particles and len(env.primitives) == n_tools
    print("case %d spheres: %d particles, %d substeps per env step" % (n_tools, n, sim.substeps), flush=True)
    gm = GradModel(env, softness=666.0, output_grid=(1,))
    obs = gm.reset()
    x0 = sim.get_state(0)[0]
    H = 2
    acts = np.random.default_rng(seed).uniform(-1, 1, (H, 3 * n_tools))
    a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
    ng = sim.n_grid
    target = np.random.default_rng(seed + 1).uniform(0, 2e-4, (ng, ng, ng))
    for s in range(H):
        obs = gm.forward(s, a[s], *obs)
    loss = (obs[2] - torch.as_tensor(target, dtype=torch.float32, device="cuda")).abs().sum()
    loss.backward()
    g = a.grad.detach().cpu().double().numpy()
    T = H * sim.substeps
    xT = sim.get_state(T)[0]

    sim_cfg = dict(cfg["SIMULATOR"])
    osim, c = oracle_sim(sim_cfg)
    assert c["substeps"] == sim.substeps
    prims = cfg["PRIMITIVES"]
    st = np.zeros((n, 24))
    st[:, 0:3] = x0
    st[:, 15:24] = np.eye(3).reshape(-1)
    res = {}
    for bits in (32, 64):
        o = Oracle(osim, [oracle_prim(p) for p in prims], bits=bits, threads=16)
        ro = OracleRollout(o, c["substeps"])
        frames, _ = ro.run(st, np.stack([prim_state0(p) for p in prims]), acts)
        m = o.grid_mass(frames[T]).astype(np.float64)
        ga, _, _ = ro.grad({T: o.grid_mass_grad(frames[T], np.sign(m - target).astype(o.dtype))})
        res[bits] = (frames[T], float(np.abs(m - target).sum()), np.asarray(ga, np.float64))
    dx = 1.0 / ng
    rel = lambda u, v: float(np.abs(u - v).max() / np.abs(v).max())  # noqa: E731
    xerr = lambda u, v: float(np.abs(u[:, 0:3] - v[:, 0:3]).max() / dx)  # noqa: E731
    print("x err vs fp32 oracle %.3g dx, vs fp64 %.3g dx (fp32 oracle vs fp64 %.3g dx); loss %.7g vs %.7g; "
          "grad rel vs fp32 oracle %.3g, vs fp64 %.3g"
          % (xerr(xT, res[32][0]), xerr(xT, res[64][0]), xerr(res[32][0], res[64][0]), float(loss), res[64][1],
             rel(g, res[32][2]), rel(g, res[64][2])))
    assert xerr(xT, res[32][0]) < 1e-4
    assert abs(float(loss) - res[64][1]) < 1e-4 * res[64][1]
    assert rel(g, res[32][2]) < 1e-3
    # every sphere moved the dough: each tool's actions carry gradient
    assert all(np.abs(g[:, 3 * j:3 * j + 3]).max() > 0 for j in range(n_tools))


def test_six_sphere_task_loads_and_matches_oracle():
    rollout_matches_oracle(triplemove_cfg(), 3 * 3333, 6, seed=7)


def test_eight_sphere_scene_matches_oracle():
    rollout_matches_oracle(eight_sphere_cfg(), 2 * 1500, 8, seed=11)


def test_too_many_collision_units_refused():
    from diffskill_amd.engine import MPMSimulator, Primitives
    grip = dict(shape="Gripper", size=(0.01, 0.03, 0.03), minimal_gap=0.02, init_gap=0.1,
                action=dict(dim=7, scale=(0.01,) * 7))
    sim_cfg = dict(n_grid=32, dt=2e-4, substeps=2, max_steps=8, n_particles=100)
    with pytest.raises(RuntimeError, match="collision units"):
        MPMSimulator(sim_cfg, Primitives([grip] * 5)).initialize(100)   # 10 jaws
    MPMSimulator(sim_cfg, Primitives([grip] * 4)).initialize(100)        # 8 jaws: accepted


def test_adjoint_entry_points_refuse_bad_frames():
    from diffskill_amd.engine import MPMSimulator, Primitives
    cap = dict(shape="Capsule", h=0.1, r=0.03, init_pos=(0.5, 0.2, 0.5), action=dict(dim=3, scale=(0.01,) * 3))
    sim_cfg = dict(n_grid=32, dt=2e-4, substeps=2, max_steps=8, n_particles=200)
    prims = Primitives([cap])
    sim = MPMSimulator(sim_cfg, prims)
    st = np.zeros((200, 24))
    st[:, 0:3] = np.random.default_rng(0).uniform(0.4, 0.6, (200, 3))
    st[:, 15:24] = np.eye(3).reshape(-1)
    sim.set_state(0, state_list(st, [prim_state0(cap)]))
    for f in range(3):
        sim.substep(f)
    e = sim.engine
    e.clear_grads(-1)
    xv = torch.zeros((e.stride, 6), device="cuda")
    ng = 32 ** 3
    grid = torch.zeros(ng, device="cuda")
    for bad in (-1, e.max_frames, e.max_frames + 5, 4):     # 4: not yet simulated
        with pytest.raises(RuntimeError, match="frame"):
            e.add_obs_grad(bad, xv)
        with pytest.raises(RuntimeError, match="frame"):
            e.grid_mass_bwd(bad, grid)
        with pytest.raises(RuntimeError, match="frame"):
            e.density_loss_bwd(bad, grid, grid)
        with pytest.raises(RuntimeError, match="frame"):
            e.min_dist_bwd(bad, torch.zeros((1, e.stride), device="cuda"))
    # the refusals left the chain without a top: frame 3 may still start it
    e.add_obs_grad(3, xv)
    with pytest.raises(RuntimeError, match="adjoint frame"):
        e.add_obs_grad(2, xv)
    e.substep_bwd(0, 3)
    torch.cuda.synchronize()
