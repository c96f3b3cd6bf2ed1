"""GPU parity of the tool-tool collision projection (set_surface_points /
set_collision_idx / apply_collision_projection, mpm_simulator.py:286-305,
and their adjoint) against the fp64 oracle: per-frame collision indices,
chosen points and projected tool poses, particle states, and action
gradients through GradModel, checkpointed recomputation and the packed
multi-scene path (eager and HIP-graph replays)."""
import numpy as np
import pytest
import torch

from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state
from scenes import oracle_sim, particles_box, prim_state0, state_list
from test_oracle_collision import rand_num

pytestmark = pytest.mark.gpu

DOUGH = ([0.4, 0.05, 0.4], [0.6, 0.12, 0.6])
SIM = dict(yield_stress=50.0, E=5e3, nu=0.2, ground_friction=1.5, gravity=(0, -9.8, 0), n_grid=32, dt=2e-4,
           substeps=3, max_steps=16, n_particles=2000)


def _q(*q):
    q = np.asarray(q, np.float64)
    return tuple(q / np.linalg.norm(q))


def box_on_obstacle():
    """A box tool pressed down onto a box obstacle that overhangs a dough block."""
    tool = dict(shape="Box", size=(0.07, 0.02, 0.05), init_pos=(0.5, 0.2, 0.5), init_rot=_q(0.95, 0.1, 0.2, 0.05),
                friction=0.9, collision_group=(0, 1), action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05)))
    obst = dict(shape="Box", size=(0.2, 0.05, 0.15), init_pos=(0.5, 0.14, 0.62), init_rot=(1.0, 0.0, 0.0, 0.0),
                friction=5.0, action=dict(dim=0, scale=()))
    acts = np.random.default_rng(3).uniform(-0.5, 0.5, (2, 6))
    acts[:, 1] = -0.8
    return [tool, obst], [(0, 1)], acts, DOUGH


def gather_like():
    """gather_move.yml's layout: a gripper and a lifter box, both colliding
    with a box obstacle (collision_group (0, 0, 1) on both tools).  The
    obstacle stands beside the dough: one gripper jaw rests on it while the
    other presses the dough."""
    grip = dict(shape="Gripper", size=(0.015, 0.04, 0.03), init_pos=(0.52, 0.16, 0.5), init_rot=_q(0.98, 0.0, 0.05, 0.1),
                init_gap=0.12, minimal_gap=0.02, friction=1.0, collision_group=(0, 0, 1),
                action=dict(dim=7, scale=(0.015, 0.01, 0.015, 0.05, 0.05, 0.1, 0.03)))
    lift = dict(shape="Box", size=(0.04, 0.02, 0.04), init_pos=(0.64, 0.17, 0.42), init_rot=_q(1.0, 0.05, 0.0, 0.1),
                friction=5.0, collision_group=(0, 0, 1), action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05)))
    obst = dict(shape="Box", size=(0.1, 0.05, 0.15), init_pos=(0.65, 0.1, 0.5), init_rot=_q(1.0, 0.0, 0.03, 0.0),
                friction=5.0, action=dict(dim=0, scale=()))
    acts = np.random.default_rng(7).uniform(-0.3, 0.3, (2, 13))
    acts[:, 1] = -0.3
    acts[:, 8] = -0.3
    return [grip, lift, obst], [(0, 2), (1, 2)], acts, ([0.3, 0.05, 0.4], [0.5, 0.12, 0.6])


def two_obstacles():
    """One tool in two pairs: projected out of the first obstacle, then out
    of the second from where the first move left it."""
    tool = dict(shape="Box", size=(0.07, 0.02, 0.05), init_pos=(0.5, 0.2, 0.5), init_rot=_q(0.95, 0.1, 0.2, 0.05),
                friction=0.9, collision_group=(0, 1, 1), action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05)))
    o1 = dict(shape="Box", size=(0.2, 0.05, 0.15), init_pos=(0.5, 0.14, 0.62), init_rot=(1.0, 0.0, 0.0, 0.0),
              friction=5.0, action=dict(dim=0, scale=()))
    o2 = dict(shape="Box", size=(0.05, 0.1, 0.05), init_pos=(0.6, 0.2, 0.45), init_rot=_q(1.0, 0.0, 0.1, 0.0),
              friction=5.0, action=dict(dim=0, scale=()))
    acts = np.random.default_rng(11).uniform(-0.5, 0.5, (2, 6))
    acts[:, 0] = 0.9
    acts[:, 1] = -0.8
    return [tool, o1, o2], [(0, 1), (0, 2)], acts, DOUGH


SCENES = {"box_on_obstacle": box_on_obstacle, "gather_like": gather_like, "two_obstacles": two_obstacles}


def dough(box, seed=5):
    return particles_box(2000, box[0], box[1], seed=seed, v_sigma=0.1)


def oracle_for(prims, pairs):
    osim, c = oracle_sim(SIM)
    o = Oracle(osim, [prim_from_cfg(p) for p in prims], collisions=(pairs, rand_num(len(pairs))))
    return o, c


@pytest.mark.parametrize("name", sorted(SCENES))
def test_collision_forward_parity(name):
    """Per frame: collision_idx equal to the oracle's, chosen points, the
    pre-move positions and every primitive state within 1e-5, particles
    within 1e-4 dx."""
    from diffskill_amd.engine import MPMSimulator, Primitives
    print("case collision forward %s" % name, flush=True)
    prims, pairs, acts, box = SCENES[name]()
    st = dough(box)
    P = Primitives(prims)
    sim = MPMSimulator(SIM, P)
    sim.set_state(0, state_list(st, [prim_state0(p) for p in prims]))
    P.set_softness(666.0)
    assert sim.collision_pairs == [list(p) for p in pairs]
    o, c = oracle_for(prims, pairs)
    S = c["substeps"]
    ro = OracleRollout(o, S)
    frames, ps = ro.run(st, np.stack([prim_state_from_cfg(p) for p in prims]), acts)
    for s in range(2):
        P.set_action(s, S, acts[s])
        for i in range(s * S, (s + 1) * S):
            sim.substep(i)
    contacts = 0
    for f in range(1, 2 * S + 1):
        idx, pts, pre = sim.get_collisions(f)
        rec = ro.colls[f - 1]
        np.testing.assert_array_equal(idx, rec["idx"], err_msg="frame %d" % f)
        contacts += int((idx >= 0).sum())
        hit = idx >= 0
        np.testing.assert_allclose(pts[hit], rec["pt"][hit], rtol=0, atol=1e-5)
        np.testing.assert_allclose(pre, rec["pre"], rtol=0, atol=1e-5)
        for j, p in enumerate(P):
            got = p.get_state(f)
            np.testing.assert_allclose(got, ps[f][j, :len(got)], rtol=0, atol=1e-5, err_msg="frame %d prim %d" % (f, j))
    assert contacts >= 2 * S, "the scene must exercise the projection"
    got = sim.get_state(2 * S)
    assert np.abs(got[0] - frames[2 * S][:, 0:3]).max() < 1e-4 * c["dx"]


def _loss_grads_gpu(prims, st, acts, target, ckpt_k=None):
    from diffskill_amd.engine import GradModel, TaichiEnv
    env = TaichiEnv(dict(SIMULATOR=SIM, PRIMITIVES=prims, SHAPES=[]), init_particles=st[:, :3], ckpt_k=ckpt_k)
    env.initialize()
    env.simulator.set_state(0, state_list(st, [prim_state0(p) for p in prims]))
    env.primitives.set_softness(666.0)
    gm = GradModel(env, softness=666.0, output_grid=(1,))
    obs = gm.reset()
    a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
    tgt = torch.as_tensor(target, dtype=torch.float32, device="cuda")
    wp = torch.as_tensor(np.random.default_rng(12).normal(size=(len(prims), 8)), dtype=torch.float32, device="cuda")
    for s in range(len(acts)):
        obs = gm.forward(s, a[s], *obs)
    # obs[1]: primitive states (n_prims, 8) of the last frame, gap 0 without one
    loss = (obs[2] - tgt).abs().sum() + 1e-3 * (obs[0] ** 2).sum() + (wp * obs[1]).sum()
    loss.backward()
    return float(loss.detach()), a.grad.detach().cpu().double().numpy(), wp.cpu().double().numpy()


def _loss_grads_oracle(prims, pairs, st, acts, target, wp):
    o, c = oracle_for(prims, pairs)
    ro = OracleRollout(o, c["substeps"])
    frames, ps = ro.run(st, np.stack([prim_state_from_cfg(p) for p in prims]), acts)
    T = len(frames) - 1
    m = o.grid_mass(frames[T])
    dims = [8 if p["shape"] in ("Gripper", "Gripper2", "Chopsticks") else 7 for p in prims]
    wpo = np.zeros((len(prims), 8))
    for j, d in enumerate(dims):
        wpo[j, :d] = wp[j, :d]
    loss = np.abs(m - target).sum() + 1e-3 * (frames[T][:, 0:6] ** 2).sum() + (wpo * ps[T]).sum()
    adj = o.grid_mass_grad(frames[T], np.sign(m - target))
    adj[:, 0:6] += 2e-3 * frames[T][:, 0:6]
    ga, _, _ = ro.grad({T: adj}, {T: wpo})
    return loss, ga


@pytest.mark.parametrize("name,ckpt_k", [("box_on_obstacle", None), ("gather_like", None), ("two_obstacles", None),
                                         ("gather_like", 2)])
def test_collision_action_grad_parity(name, ckpt_k):
    """Action gradients of a mass-grid L1 + state loss + a loss on the final
    tool poses, through GradModel, within 1e-3 of the fp64 oracle (with
    checkpointed recomputation re-running the projection)."""
    print("case collision grads %s ckpt_k=%s" % (name, ckpt_k), flush=True)
    prims, pairs, acts, box = SCENES[name]()
    st = dough(box)
    target = np.random.default_rng(2).uniform(0, 1e-4, (32, 32, 32))
    loss, g, wp = _loss_grads_gpu(prims, st, acts, target, ckpt_k=ckpt_k)
    loss_ref, g_ref = _loss_grads_oracle(prims, pairs, st, acts, target, wp)
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print("%s: loss %.8g vs %.8g, grad rel %.3g" % (name, loss, loss_ref, err))
    assert abs(loss - loss_ref) < 1e-5 * abs(loss_ref)
    assert err < 1e-3, (g, g_ref)


def test_collision_in_packed_scenes_eager_and_graph():
    """Two scenes with collision pairs packed in one context: per-scene
    action gradients of the batched optimiser within 1e-3 of the oracle, and
    HIP-graph replays equal to the eager pass."""
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    print("case collision packed scenes", flush=True)
    H, S = 2, 3
    specs, oracle_cases = [], []
    for k, name in enumerate(("box_on_obstacle", "gather_like")):
        prims, pairs, acts, box = SCENES[name]()
        st = dough(box, seed=20 + k)
        specs.append(dict(sim=SIM, prims=prims, x0=st[:, :3]))
        oracle_cases.append((prims, pairs, acts, reset_state(st[:, :3])))
    b = SceneBatch(specs, H * S + 1)
    P = b.n_prims
    table = np.zeros((H, P, 7), np.float32)
    o0 = 0
    for prims, _, acts, _ in oracle_cases:
        c0 = 0
        for j, p in enumerate(prims):
            d = p["action"]["dim"]
            table[:, o0 + j, :d] = acts[:, c0:c0 + d]
            c0 += d
        o0 += len(prims)
    targets = torch.as_tensor(np.random.default_rng(4).uniform(0, 1e-4, (2, 32, 32, 32)), dtype=torch.float32,
                              device="cuda")
    opt = BatchedTrajOpt(b, H, S, targets, lr=0.0, init_actions=torch.as_tensor(table), use_graph=False)
    opt._device_pass()
    torch.cuda.synchronize()
    g = opt.grads.detach().cpu().double().numpy()
    losses = (opt.mass - targets).abs().sum(dim=(1, 2, 3)).cpu().double().numpy()
    o0 = 0
    for k, (prims, pairs, acts, st0) in enumerate(oracle_cases):
        o, c = oracle_for(prims, pairs)
        ro = OracleRollout(o, S)
        frames, _ = ro.run(st0, np.stack([prim_state_from_cfg(p) for p in prims]), acts)
        T = len(frames) - 1
        m = o.grid_mass(frames[T])
        tg = targets[k].cpu().double().numpy()
        ga, _, _ = ro.grad({T: o.grid_mass_grad(frames[T], np.sign(m - tg))})
        gk = np.concatenate([g[:, o0 + j, :p["action"]["dim"]] for j, p in enumerate(prims)], axis=1)
        err = np.abs(gk - ga).max() / np.abs(ga).max()
        print("scene %d: loss %.8g vs %.8g, grad rel %.3g" % (k, losses[k], np.abs(m - tg).sum(), err))
        assert abs(losses[k] - np.abs(m - tg).sum()) <= 1e-4 * np.abs(m - tg).sum()
        assert err < 1e-3, (k, gk, ga)
        o0 += len(prims)
    # graph replays against the eager pass (float atomics: run-to-run noise only)
    opt_g = BatchedTrajOpt(b, H, S, targets, lr=0.0, init_actions=torch.as_tensor(table), use_graph=True)
    for r in range(3):
        opt_g.iteration()
        torch.cuda.synchronize()
        gg = opt_g.grads.detach().cpu().double().numpy()
        assert np.abs(gg - g).max() <= 1e-4 * np.abs(g).max(), r
        idx_g = [b.engine.get_collisions(f)[0] for f in range(1, H * S + 1)]
        assert all((i >= 0).any() for i in idx_g)
