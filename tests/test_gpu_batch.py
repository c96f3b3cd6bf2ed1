"""GPU tests of the packed multi-scene path (SceneBatch / BatchedTrajOpt):
gradients against the fp64 oracle per scene, packed against one context per
scene, adjoints that reuse the forward's binning and grid against adjoints
that rebuild them (including a kept-grid capacity too small for the pass), and
HIP-graph replays against the eager pass."""
import numpy as np
import pytest
import torch

from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg

pytestmark = pytest.mark.gpu


def small_workload(n_scenes=3, H=2, S=5):
    from diffskill_amd import workloads
    return workloads.config4(n_scenes=n_scenes, n_env_steps=H, substeps=S, lo=1000, hi=3000)


def targets_for(w):
    from diffskill_amd.batch import SceneBatch
    tb = SceneBatch([dict(sim=s["sim"], prims=[], x0=t) for s, t in zip(w["scenes"], w["target_x"])], 2)
    out = tb.grid_mass(0).clone()
    torch.cuda.synchronize()
    return out


def run_pass(w, targets, ckpt_k=0, keep_blocks=-1, graph=False):
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    H, S = w["H"], w["substeps"]
    b = SceneBatch(w["scenes"], H * S + 1, ckpt_k=ckpt_k, keep_blocks=keep_blocks)
    init = torch.as_tensor(workloads.action_tensor(w["actions"], H))
    opt = BatchedTrajOpt(b, H, S, targets, lr=0.0, init_actions=init, use_graph=graph)
    opt._device_pass()
    torch.cuda.synchronize()
    return opt, opt.action_grads().detach().cpu().double().numpy(), opt.losses().detach().cpu().double().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def test_batched_grads_match_oracle_per_scene():
    w = small_workload()
    targets = targets_for(w)
    opt, g, losses = run_pass(w, targets)
    H, S = w["H"], w["substeps"]
    init = np.clip(np.concatenate(w["actions"], axis=1), -1, 1)  # (H, P, 3)
    o0 = 0
    for k, sc in enumerate(w["scenes"]):
        print("case batch-vs-oracle scene %d (%d particles)" % (k, len(sc["x0"])), flush=True)
        npk = len(sc["prims"])
        osim, _ = sim_from_cfg(sc["sim"])
        o = Oracle(osim, [prim_from_cfg(p) for p in sc["prims"]], bits=64, threads=4)
        ro = OracleRollout(o, S)
        acts = np.concatenate([init[:, o0 + j, :3] for j in range(npk)], axis=1)
        frames, _ = ro.run(reset_state(sc["x0"]), np.stack([prim_state_from_cfg(p) for p in sc["prims"]]), acts)
        T = len(frames) - 1
        m = o.grid_mass(frames[T])
        tg = targets[k].cpu().double().numpy()
        loss_ref = np.abs(m - tg).sum()
        ga, _, _ = ro.grad({T: o.grid_mass_grad(frames[T], np.sign(m - tg))})
        gk = np.concatenate([g[:, o0 + j, :3] for j in range(npk)], axis=1)
        assert abs(losses[k] - loss_ref) <= 1e-4 * abs(loss_ref)
        assert rel(gk, ga) <= 1e-3, (k, rel(gk, ga))
        o0 += npk


def test_fused_density_loss_adjoint_matches_grid_adjoint():
    """mpm_density_loss_bwd (node adjoint weight * sign(m - target) formed in
    the grid-mass adjoint) against mpm_grid_mass_bwd fed that sign grid:
    bit-exact, including nodes where mass equals target (sign 0) and a
    weight other than 1; a grid of the wrong size is refused on the host."""
    from diffskill_amd import workloads
    from diffskill_amd.batch import SceneBatch
    w = small_workload()
    H, S = w["H"], w["substeps"]
    T = H * S
    b = SceneBatch(w["scenes"], T + 1)
    e = b.engine
    b.restart()
    table = torch.as_tensor(workloads.action_tensor(w["actions"], H)).cuda().clamp(-1, 1)
    for s in range(H):
        e.set_actions(s, S, table[s])
    e.substep_fwd(0, T)
    n = b.consts[0]["n_grid"]
    mass = torch.zeros((b.n_scenes, n, n, n), dtype=torch.float32, device="cuda")
    e.grid_mass(T, mass)
    target = targets_for(w)
    target[0, :8] = mass[0, :8]          # equal nodes, inside and outside the clouds
    for weight in (1.0, 0.37):
        print("case fused density adjoint weight %g" % weight, flush=True)
        g = torch.sign(mass - target) * weight
        e.clear_grads(T)
        e.grid_mass_bwd(T, g)
        ref = e.adjoint(T).clone()
        e.clear_grads(T)
        e.density_loss_bwd(T, mass, target, weight)
        got = e.adjoint(T).clone()
        torch.cuda.synchronize()
        assert torch.count_nonzero(ref) > 0
        assert torch.equal(got, ref)
    with pytest.raises(ValueError):
        e.density_loss_bwd(T, mass[:1], target)


@pytest.mark.parametrize("ckpt_k,keep_blocks", [(0, 0), (0, 1), (3, -1), (3, 0)])
def test_kept_and_rebuilt_adjoint_grids_agree(ckpt_k, keep_blocks):
    """keep_blocks=-1 (default: the adjoint reads the forward's binning and
    grid), 0 (rebuilds both every substep) and 1 (capacity overflows: the
    device falls back to rebuilding) give the same gradients, store-all and
    checkpointed."""
    print("case kept-grid ckpt_k=%d keep_blocks=%d" % (ckpt_k, keep_blocks), flush=True)
    w = small_workload()
    targets = targets_for(w)
    _, g_ref, l_ref = run_pass(w, targets, ckpt_k=0, keep_blocks=-1)
    _, g, l = run_pass(w, targets, ckpt_k=ckpt_k, keep_blocks=keep_blocks)
    np.testing.assert_allclose(l, l_ref, rtol=1e-5)
    assert rel(g, g_ref) <= 1e-4


def test_overflow_fallback_under_graph_replay():
    """keep_blocks=1: every substep's kept grid overflows, so each substep's
    grid is rebuilt for the adjoint by workgroups riding on the previous
    adjoint substep's g2p and grid-adjoint launches.  Three replays of the
    captured iteration (eager work between them) each give the gradients of
    the eager kept-grid pass."""
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    print("case overflow fallback under graph replay", flush=True)
    w = small_workload()
    targets = targets_for(w)
    _, g_ref, _ = run_pass(w, targets)
    H, S = w["H"], w["substeps"]
    b = SceneBatch(w["scenes"], H * S + 1, keep_blocks=1)
    init = torch.as_tensor(workloads.action_tensor(w["actions"], H))
    opt = BatchedTrajOpt(b, H, S, targets, lr=0.0, init_actions=init, use_graph=True)
    for _ in range(3):
        opt.iteration()
        torch.cuda.synchronize()
        g = opt.action_grads().detach().cpu().double().numpy()
        assert rel(g, g_ref) <= 1e-4
        _ = torch.ones(1 << 20, device="cuda").sum().item()


def test_batched_matches_per_scene_contexts():
    from diffskill_amd import workloads
    from diffskill_amd.batch import PlainTrajOpt
    w = small_workload()
    targets = targets_for(w)
    _, g, l = run_pass(w, targets)
    init = torch.as_tensor(workloads.action_tensor(w["actions"], w["H"]))
    p = PlainTrajOpt(w["scenes"], w["H"], w["substeps"], targets, lr=0.0, init_actions=init)
    gp = []
    for part in p.parts:
        part._device_pass()
        gp.append(part.action_grads()[:, :part.b.n_prims])
    torch.cuda.synchronize()
    gp = torch.cat(gp, dim=1).detach().cpu().double().numpy()
    np.testing.assert_allclose(p.losses().detach().cpu().double().numpy(), l, rtol=1e-5)
    assert rel(gp, g[:, :gp.shape[1]]) <= 1e-4


def test_empty_scene_in_a_batch():
    """A scene without particles packed between two others: the others'
    losses and action gradients are those of the batch without it (to the
    float-atomics order noise), its own loss is its target's total mass and
    its action gradients are 0."""
    w = small_workload()
    targets = targets_for(w)
    _, g, l = run_pass(w, targets)
    empty = dict(w["scenes"][1], x0=np.zeros((0, 3)))
    w2 = dict(w, scenes=w["scenes"][:1] + [empty] + w["scenes"][1:], actions=w["actions"][:1] + [w["actions"][1]] + w["actions"][1:])
    t2 = torch.cat([targets[:1], targets[1:2], targets[1:]])
    print("case empty scene in a batch of %d" % len(w2["scenes"]), flush=True)
    _, g2, l2 = run_pass(w2, t2)
    p = [a.shape[1] for a in w["actions"]]
    np.testing.assert_allclose(l2[1], float(targets[1].double().sum()), rtol=1e-6)
    np.testing.assert_allclose(np.delete(l2, 1), l, rtol=1e-5)
    assert np.all(g2[:, p[0]:p[0] + p[1]] == 0)
    keep = np.r_[0:p[0], p[0] + p[1]:g2.shape[1]]
    assert rel(g2[:, keep], g[:, :len(keep)]) <= 1e-4


def test_graph_replays_match_eager():
    """The captured forward + adjoint pass, replayed several times with eager
    work and synchronisation in between, reproduces the eager result."""
    w = small_workload()
    targets = targets_for(w)
    opt, g_eager, l_eager = run_pass(w, targets, graph=True)
    for r in range(3):
        opt.iteration()   # lr = 0: the actions stay put
        torch.cuda.synchronize()
        g = opt.action_grads().detach().cpu().double().numpy()
        l = opt.losses().detach().cpu().double().numpy()
        print("case graph replay %d: rel %.3g" % (r, rel(g, g_eager)), flush=True)
        np.testing.assert_allclose(l, l_eager, rtol=1e-5)
        assert rel(g, g_eager) <= 1e-4
        opt._device_pass()   # eager work between replays
        torch.cuda.synchronize()


def test_graph_adam_iterations_match_eager():
    """The captured iteration includes the Adam step (capturable Adam, state
    created up front).  Each of three replays moves the actions exactly as
    torch's eager Adam (float64, on the CPU) does when fed the gradient that
    replay's pass produced; an eager pass of another optimiser runs between the
    replays.  (Graph and eager passes agree to float-atomics rounding, checked
    above; Adam's first steps lr * g / (|g| + eps) amplify that rounding for
    near-zero components, so the step is checked on the replay's own gradient.)"""
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    w = small_workload()
    targets = targets_for(w)
    H, S = w["H"], w["substeps"]
    init = torch.as_tensor(workloads.action_tensor(w["actions"], H))
    b = SceneBatch(w["scenes"], H * S + 1)
    og = BatchedTrajOpt(b, H, S, targets, lr=1e-2, init_actions=init, use_graph=True)
    oe = BatchedTrajOpt(b, H, S, targets, lr=1e-2, init_actions=init, use_graph=False)
    ref = torch.nn.Parameter(init.to(torch.float64).clone())
    ref_opt = torch.optim.Adam([ref], lr=1e-2)
    for r in range(3):
        og.iteration()
        torch.cuda.synchronize()
        ref.grad = og.action_grads().detach().cpu().double()
        ref_opt.step()
        ag = og.actions.detach().cpu().double()
        d = (ag - ref.detach()).abs().max().item()
        moved = (ref.detach() - init.double()).abs().max().item()
        print("case graph adam %d: moved %.3g, |graph - torch Adam| %.3g" % (r, moved, d), flush=True)
        assert moved > 1e-3 and d <= 1e-6
        oe._device_pass()   # eager work between the replays
        torch.cuda.synchronize()


def test_launch_shapes_agree_and_validate():
    """Particles per work item and workgroup counts change scheduling only:
    every shape gives the same gradients to rounding; invalid shapes raise."""
    w = small_workload()
    targets = targets_for(w)
    opt, g_ref, l_ref = run_pass(w, targets)
    e = opt.b.engine
    for shape in ((256, 64, 64), (512, 1024, 2048), (2048, 16, 32), (8192, 0, 0)):
        print("case launch shape %s" % (shape,), flush=True)
        e.set_launch(*shape)
        opt._device_pass()
        torch.cuda.synchronize()
        g = opt.action_grads().detach().cpu().double().numpy()
        np.testing.assert_allclose(opt.losses().detach().cpu().double().numpy(), l_ref, rtol=1e-5)
        assert rel(g, g_ref) <= 1e-4
    assert e.get_launch()["particles_per_item"] == 8192
    for bad in ((100, 0, 0), (16384, 0, 0), (0, -1, 0)):
        with pytest.raises(RuntimeError):
            e.set_launch(*bad)
