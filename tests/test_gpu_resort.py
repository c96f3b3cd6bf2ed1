"""The engine's slot re-sort (mpm_set_resort): every R substeps the particles'
slots are relabelled into the cell order of the binning, the adjoint maps
back across each boundary and a checkpointed recompute applies the stored
orders.  Losses, mass grids and action gradients do not depend on slot
order: with the re-sort they equal the unsorted run up to float summation
order, bitwise across repeats, checkpointing and graph replays in the
fixed-order mode, and each particle's state (found through mpm_slot_order)
is the unsorted run's."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def workload():
    from diffskill_amd import workloads
    return workloads.config4(n_scenes=4, n_env_steps=4, substeps=10, lo=1500, hi=6000)


def targets_for(w):
    from diffskill_amd.batch import SceneBatch
    tb = SceneBatch([dict(sim=s["sim"], prims=[], x0=t) for s, t in zip(w["scenes"], w["target_x"])], 2)
    out = tb.grid_mass(0).clone()
    torch.cuda.synchronize()
    return out


def run(w, targets, resort=True, det=True, ckpt_k=0, graph=False, reps=1):
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    H, S = w["H"], w["substeps"]
    b = SceneBatch(w["scenes"], H * S + 1, ckpt_k=ckpt_k, deterministic=det)
    init = torch.as_tensor(workloads.action_tensor(w["actions"], H))
    opt = BatchedTrajOpt(b, H, S, targets, lr=0.0, init_actions=init, use_graph=graph, resort=resort)
    assert b.engine.resort == (S if resort else 0)
    outs = []
    for _ in range(reps):
        if graph:
            opt.iteration()
        else:
            opt._device_pass()
        torch.cuda.synchronize()
        T = H * S
        xT = b.engine.frame(T)[:3].detach().clone()
        if resort:
            # the row of frame 0's slot c: its particle's slot in frame T
            xT = xT[:, b.engine.slot_order(T)]
        outs.append(dict(g=opt.action_grads().detach().cpu().numpy().copy(), loss=opt.losses().detach().cpu().numpy().copy(),
                         mass=opt.mass.detach().cpu().numpy().copy(), xT=xT.cpu().numpy().copy()))
        _ = torch.ones(1 << 20, device="cuda").sum().item()   # eager work between passes
    return outs, b


def rel(a, b):
    return float(np.abs(a.astype(np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("g", "loss", "mass", "xT"))


def test_resort_matches_the_unsorted_run():
    w = workload()
    targets = targets_for(w)
    print("case resort vs unsorted", flush=True)
    ref, _ = run(w, targets, resort=False)
    out, b = run(w, targets, resort=True)
    n = int(sum(len(s["x0"]) for s in w["scenes"]))
    print("grad rel %.3g, loss rel %.3g, mass rel %.3g" % (rel(out[0]["g"], ref[0]["g"]), rel(out[0]["loss"], ref[0]["loss"]),
                                                          rel(out[0]["mass"], ref[0]["mass"])))
    assert rel(out[0]["g"], ref[0]["g"]) <= 1e-4
    assert rel(out[0]["loss"], ref[0]["loss"]) <= 1e-5
    assert rel(out[0]["mass"], ref[0]["mass"]) <= 1e-5
    # every particle ends where the unsorted run put it (slots of frame 0 are
    # the same in both runs); the padding columns between scenes hold nothing
    cols = np.concatenate([np.arange(b.engine.offsets[k], b.engine.offsets[k] + len(s["x0"])) for k, s in enumerate(w["scenes"])])
    assert len(cols) == n
    dx = np.abs(out[0]["xT"][:, cols] - ref[0]["xT"][:, cols]).max()
    print("max |x_T difference| %.3g" % dx)
    assert dx <= 1e-5


def test_resort_is_bitwise_across_repeats_checkpoints_and_graph_replays():
    w = workload()
    targets = targets_for(w)
    print("case resort repeats", flush=True)
    ref, _ = run(w, targets, reps=2)
    assert same(ref[0], ref[1])
    for K in (3, 7):
        print("case resort ckpt_k %d" % K, flush=True)
        out, _ = run(w, targets, ckpt_k=K)
        assert same(out[0], ref[0]), K
    print("case resort graph replays", flush=True)
    out, _ = run(w, targets, graph=True, reps=3)
    for o in out:
        assert same(o, ref[0])


def test_resort_in_the_unordered_mode():
    w = workload()
    targets = targets_for(w)
    print("case resort unordered mode", flush=True)
    ref, _ = run(w, targets, resort=False, det=False)
    out, _ = run(w, targets, resort=True, det=False, ckpt_k=7)
    assert rel(out[0]["g"], ref[0]["g"]) <= 1e-4
    assert rel(out[0]["mass"], ref[0]["mass"]) <= 1e-5


def test_resort_refuses_copies_across_slot_orders():
    w = workload()
    targets = targets_for(w)
    _, b = run(w, targets)
    T = w["H"] * w["substeps"]
    with pytest.raises(RuntimeError, match="slot orders"):
        b.engine.copy_frame(T, 0)
    b.engine.copy_frame(1, 2)   # same slot order: allowed


def test_resort_arguments_are_checked():
    """Intervals below 8 substeps and buffers smaller than mpm_resort_bytes
    are refused; turning the re-sort off needs no buffer."""
    import ctypes
    from diffskill_amd.batch import SceneBatch
    w = workload()
    b = SceneBatch(w["scenes"], w["H"] * w["substeps"] + 1)
    e = b.engine
    with pytest.raises(RuntimeError, match="at least 8"):
        e.set_resort(4)
    need = e.L.mpm_resort_bytes(e.ctx, 10)
    assert need > 0
    small = torch.empty(int(need) - 256, dtype=torch.uint8, device="cuda")
    assert e.L.mpm_set_resort(e.ctx, 10, ctypes.c_void_p(small.data_ptr()), int(small.numel())) != 0
    assert e.resort == 0
    e.set_resort(10)
    assert e.resort == 10
    e.set_resort(0)
    assert e.resort == 0
