#!/usr/bin/env python
"""Benchmark: fwd+adjoint particle-substeps/s of batched differentiable MLS-MPM.

Workload (one "step" = one trajectory-optimisation iteration): 64 scenes,
particle counts log-uniform in [2k, 40k], each with its own 64^3 grid and 1-2
capsule/box tools; 200 substeps forward + 200 adjoint substeps over all scenes,
mass-grid L1 loss, action adjoints, one Adam step (lr 1e-2).  Inputs are
resident in HBM before timing.  Reported next to it:
- the gradient accuracy gate: the batched path's action gradients on four
  scenes (the two smallest, the smallest other capsule+box scene, the
  largest) against the fp32 CPU oracle (the ti.f32 arithmetic of the
  reference; the gate, <= 1e-3 relative) and the fp64 one (stated);
- the plain per-scene GPU path (one context per scene, same kernels, no
  packing, no graph);
- forward-only throughput of the same workload, and of the plain forward
  dough block (5k particles, 500 substeps);
- the gather scene's device memory, store-all against checkpoints every K
  in {10, 50, 100} frames, and the large cutting scene's fwd+adjoint rate;
- live HIP-event kernel timings with the HBM roofline of the dominant kernel;
- the fp32 CPU oracle timed on a bounded sample on all host cores the
  machine grants this job and on one core.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "fwd+adjoint particle-substeps/s, batched vs plain, at grad rel err ≤1e-3"
HBM_PEAK_GBS = 8000.0
GRAD_GATE = 1e-3
HOST_CORES = 16   # the host-core share of one GPU on the measurement box

# algorithmic bytes per particle for each particle kernel: the particle-state
# (and particle-adjoint) floats it must read and write once, fp32
ALGO_BYTES = {
    "stress": 72 + 72,               # read C F, write F' and the affine matrix
    "p2g": 60,                       # read x v and the affine matrix
    "g2p": 12 + 60,                  # read x, write x' v' C'
    "p2g_grid": 60,                  # read x v and the affine matrix
    "g2p_bwd": 12 + 12 + 60 + 12,    # read x, x', adj x' v' C', write adj x
    "p2g_bwd": 60 + 12 + 24 + 36,    # read x v affine, adj x; write adj x v and the affine adjoint
    "stress_bwd": 72 + 36 + 36 + 72,  # read C F, affine adjoint, adj F'; write adj C F
    "bin": 12,
}


# engine phase -> kernel name in the rocprof summaries
PHASE_KERNEL = {"stress": "k_stress", "p2g": "k_p2g_scatter", "p2g_grid": "k_p2g_scatter_rebuild", "g2p": "k_g2p_tile",
                "g2p_bwd": "k_g2p_bwd_tile", "p2g_bwd": "k_p2g_bwd_gather", "stress_bwd": "k_stress_bwd", "bin": "k_bin",
                "grid_op": "k_grid_op", "grid_op_keep": "k_grid_op_rebuild", "grid_op_bwd": "k_grid_op_bwd"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_batched64_config4.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_issue_batched64_config4.json")


def measured_traffic(phase):
    """HBM bytes per launch of the phase's kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)[PHASE_KERNEL[phase]]
        return int(t["read_bytes"] + t["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def measured_valu(phase):
    """(VALU wave-instructions per launch of the phase's kernel, issue peak per
    second) from the committed rocprofv3 SQ counter summary, or None."""
    try:
        with open(VALU_FILE) as f:
            t = json.load(f)
        return t["kernels"][PHASE_KERNEL[phase]]["valu_instr_per_dispatch"], t["issue_peak_per_s"]
    except (OSError, KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_targets(w):
    """Target mass grids: compute_grid_m of the slab target particles (same engine)."""
    import torch
    from diffskill_amd.batch import SceneBatch
    tscenes = [dict(sim=s["sim"], prims=[], x0=t) for s, t in zip(w["scenes"], w["target_x"])]
    tb = SceneBatch(tscenes, 2)
    out = tb.grid_mass(0)
    torch.cuda.synchronize()
    del tb
    return out


def check_scene_ids(sizes, scenes):
    """The two smallest scenes, the smallest other scene with a box tool, the largest."""
    import numpy as np
    order = [int(k) for k in np.argsort(sizes)]
    ids = order[:2]
    box = [k for k in order if k not in ids and any(p["shape"] == "Box" for p in scenes[k]["prims"])]
    if box:
        ids.append(box[0])
    if order[-1] not in ids:
        ids.append(order[-1])
    return ids


def oracle_scene_grads(w, k, table, target_grid, bits, threads):
    """CPU oracle (fp64 or fp32): loss and action adjoint of scene k under the given action table."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg
    sc = w["scenes"][k]
    osim, S = sim_from_cfg(sc["sim"])
    o = Oracle(osim, [prim_from_cfg(p) for p in sc["prims"]], bits=bits, threads=threads)
    ro = OracleRollout(o, w["substeps"])
    dims = [p["action"]["dim"] for p in sc["prims"]]
    acts = np.concatenate([table[:, j, :d] for j, d in enumerate(dims)], axis=1)
    frames, _ = ro.run(reset_state(sc["x0"]), np.stack([prim_state_from_cfg(p) for p in sc["prims"]]), acts)
    T = len(frames) - 1
    m = o.grid_mass(frames[T]).astype(np.float64)
    tg = target_grid.astype(np.float64)
    loss = np.abs(m - tg).sum()
    ga, _, _ = ro.grad({T: o.grid_mass_grad(frames[T], np.sign(m - tg).astype(o.dtype))})
    return loss, np.asarray(ga, np.float64), dims


def rel_errs(g, ref):
    import numpy as np
    return (float(np.abs(g - ref).max() / max(np.abs(ref).max(), 1e-30)),
            float(np.linalg.norm(g - ref) / max(np.linalg.norm(ref), 1e-30)))


def cpu_baseline(w, budget_s, threads, scenes=None):
    """fp32 oracle (C + OpenMP) on whole scenes of the same workload, fwd + adjoint
    over all 200 substeps, until ~budget_s seconds of CPU work have run."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg
    done_ps, spent, used = 0, 0.0, []
    order = list(range(len(w["scenes"]))) if scenes is None else scenes
    for k in order:
        sc = w["scenes"][k]
        osim, _ = sim_from_cfg(sc["sim"])
        o = Oracle(osim, [prim_from_cfg(p) for p in sc["prims"]], bits=32, threads=threads)
        ro = OracleRollout(o, w["substeps"])
        acts = np.concatenate([w["actions"][k][:, j, :] for j in range(len(sc["prims"]))], axis=1)
        t0 = time.perf_counter()
        frames, _ = ro.run(reset_state(sc["x0"]), np.stack([prim_state_from_cfg(p) for p in sc["prims"]]), acts)
        T = len(frames) - 1
        m = o.grid_mass(frames[T])
        ro.grad({T: o.grid_mass_grad(frames[T], np.sign(m))})
        spent += time.perf_counter() - t0
        done_ps += len(sc["x0"]) * T
        used.append(k)
        if spent >= budget_s:
            break
    return done_ps / spent, used, spent


def timed(fn, reps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1000 / reps


def config1_forward(reps=5):
    """Plain forward dough block: 5k particles, capsule, 50 env steps x 10
    substeps at dt 2e-4 (the largest stable step), captured as one graph."""
    import torch
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    w = workloads.config1(dt=2e-4)
    H, S = w["H"], w["substeps"]
    b = SceneBatch([dict(sim=w["sim"], prims=w["prims"], x0=w["x0"])], H * S + 1, keep_blocks=0)
    n = b.consts[0]["n_grid"]
    init = torch.zeros((H, 1, 7))
    init[:, 0, :3] = torch.as_tensor(w["actions"], dtype=torch.float32)
    opt = BatchedTrajOpt(b, H, S, torch.zeros((1, n, n, n), device="cuda"), lr=0.0, init_actions=init)
    opt.forward()
    ms = timed(opt.forward, reps)
    return dict(workload="plain forward: 5k-particle dough block, capsule, 500 substeps at dt 2e-4, 64^3 grid",
                particles=len(w["x0"]), substeps=H * S, ms_per_pass=round(ms, 3), value=len(w["x0"]) * H * S / (ms / 1000),
                unit="particle-substeps/s (forward)")


def config3_memory():
    """Gather scene (30k particles, gripper + spatula, 1000 substeps, one-way
    Chamfer loss) through the drop-in GradModel: peak device memory and
    fwd+adjoint time for store-all and checkpoints every K frames."""
    import numpy as np
    import torch
    from diffskill_amd import workloads
    from diffskill_amd.engine import GradModel, TaichiEnv
    from diffskill_amd.pointcloud import chamfer_distance
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    w = workloads.config3()
    H = w["H"]
    acts = np.concatenate(w["actions"], axis=1)
    t = torch.as_tensor(w["target_points"], dtype=torch.float32, device="cuda")
    out = {}
    import gc
    for K in (None, 10, 50, 100):
        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        torch.cuda.reset_peak_memory_stats()
        base = torch.cuda.memory_allocated()
        env = TaichiEnv(dict(SIMULATOR=w["sim"], PRIMITIVES=w["prims"], SHAPES=[]), init_particles=w["x0"], ckpt_k=K)
        env.initialize()
        gm = GradModel(env, softness=666.0)
        # two passes, the second timed (the first includes first-use costs)
        for rep in range(2):
            t0 = time.perf_counter()
            obs = gm.reset()
            a = torch.tensor(acts, dtype=torch.float32, device="cuda", requires_grad=True)
            for s in range(H):
                obs = gm.forward(s, a[s], *obs)
            loss = chamfer_distance(obs[0][:, :3], t)
            loss.backward()
            torch.cuda.synchronize()
            sec = time.perf_counter() - t0
        peak = torch.cuda.max_memory_allocated() - base
        ws = env.simulator.engine.ws.numel()
        out["store-all" if K is None else "K=%d" % K] = dict(peak_bytes=int(peak), workspace_bytes=int(ws), seconds=round(sec, 3),
                                                            value=len(w["x0"]) * H * w["substeps"] / sec)
        del env, gm, obs, a, loss
    return dict(workload="gather scene: 30k particles, gripper + spatula, 1000 substeps fwd+adjoint through GradModel, "
                         "one-way Chamfer loss", unit="particle-substeps/s (fwd+adjoint, eager, per-env-step host calls)",
                runs=out)


def config5_rate():
    """The large cutting scene (1M particles, 128^3 grid, 2000 substeps, a
    checkpoint every 50 frames): one fwd+adjoint pass, timed."""
    import torch
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, SceneBatch
    w = workloads.config5()
    H, S = w["H"], w["substeps"]
    n = 128
    b = SceneBatch([dict(sim=w["sim"], prims=w["prims"], x0=w["x0"])], H * S + 1, ckpt_k=w["ckpt_k"])
    init = torch.zeros((H, 1, 7))
    init[:, 0, :3] = torch.as_tensor(w["actions"], dtype=torch.float32)
    opt = BatchedTrajOpt(b, H, S, torch.zeros((1, n, n, n), device="cuda"), lr=0.0, init_actions=init, use_graph=False)
    opt._device_pass()
    ms = timed(opt._device_pass, 1)
    return dict(workload="cutting scene: 1M particles, knife, 128^3 grid, 2000 substeps fwd+adjoint, checkpoint every 50",
                particles=len(w["x0"]), substeps=H * S, ms_per_pass=round(ms, 1), value=len(w["x0"]) * H * S / (ms / 1000),
                unit="particle-substeps/s (fwd+adjoint)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--ckpt-k", type=int, default=0, help="checkpoint every K frames (0 = store all)")
    ap.add_argument("--keep-blocks", type=int, default=-1,
                    help="node blocks of each substep's grid kept for the adjoint (-1 default, 0 rebuild every grid)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--resort", type=int, default=-1,
                    help="slot re-sort interval in substeps (-1: every env step, 0: off)")
    ap.add_argument("--unordered", action="store_true",
                    help="unordered float atomics instead of the engine's default fixed-order (bitwise-reproducible) mode")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--launch", default=None,
                    help="PPI,TILE,GRID: particles per work item and workgroup counts (skips the autotuner)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other workloads (configs 1, 3, 5)")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--check-scenes", default="auto",
                    help="scenes checked against the oracles: 'auto' (two smallest, a box scene, the largest) or a list")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from diffskill_amd import workloads
    from diffskill_amd.batch import BatchedTrajOpt, PlainTrajOpt, SceneBatch

    torch.cuda.init()
    # calibration kernel of known traffic for the HBM byte counters: reads and
    # writes 256 MiB (elementwise multiply), three launches
    cal = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    for _ in range(3):
        cal = cal.mul(1.0)
    torch.cuda.synchronize()
    del cal
    w = workloads.config4(n_scenes=args.scenes)
    H, S = w["H"], w["substeps"]
    T = H * S
    log("workload: %d scenes, %d particles, %d substeps" % (len(w["scenes"]), int(np.sum(w["sizes"])), T))
    batch = SceneBatch(w["scenes"], T + 1, ckpt_k=args.ckpt_k, keep_blocks=args.keep_blocks, deterministic=not args.unordered)
    targets = build_targets(w)
    init = torch.as_tensor(workloads.action_tensor(w["actions"], H))
    resort = True if args.resort < 0 else args.resort
    opt = BatchedTrajOpt(batch, H, S, targets, lr=1e-2, init_actions=init, use_graph=not args.no_graph, resort=resort)
    resort_used = batch.engine.resort
    n_total = batch.total_particles
    launch = None
    if args.launch:
        batch.engine.set_launch(*[int(v) for v in args.launch.split(",")])
    elif not args.no_tune:
        launch, tune_ms = opt.autotune()
        log("autotuned launch shape %s (%d candidates, %.1f-%.1f ms per eager pass)"
            % (launch, len(tune_ms), min(tune_ms.values()), max(tune_ms.values())))

    # ---- gradient accuracy gate: the batched path vs the fp32 oracle (the
    # reference's ti.f32 arithmetic), with the fp64 distance stated
    grad_checks, gate = [], None
    if not args.no_oracle:
        opt._device_pass()
        torch.cuda.synchronize()
        g_all = opt.action_grads().detach().cpu().double().numpy()
        losses = opt.losses().detach().cpu().numpy()
        ids = check_scene_ids(w["sizes"], w["scenes"]) if args.check_scenes == "auto" else \
            [int(v) for v in args.check_scenes.split(",") if v]
        for k in ids:
            o0 = int(batch.prim_offsets[k])
            npk = len(w["scenes"][k]["prims"])
            table = torch.clamp(init, -1, 1).numpy()[:, o0:o0 + npk]
            tgt = targets[k].cpu().numpy()
            c = dict(scene=k, particles=int(w["sizes"][k]), tools=[p["shape"] for p in w["scenes"][k]["prims"]],
                     loss=float(losses[k]))
            for bits in (32, 64):
                t0 = time.perf_counter()
                loss_ref, ga, dims = oracle_scene_grads(w, k, table, tgt, bits, HOST_CORES)
                g_flat = np.concatenate([g_all[:, o0 + j, :d] for j, d in enumerate(dims)], axis=1)
                e_max, e_l2 = rel_errs(g_flat, ga)
                c["loss_fp%d" % bits] = float(loss_ref)
                c["rel_err_max_fp%d" % bits] = e_max
                c["rel_err_l2_fp%d" % bits] = e_l2
                if bits == 32:
                    g32 = ga
                else:
                    c["fp32_oracle_vs_fp64_max"] = rel_errs(g32, ga)[0]
                log("grad check scene %d (%d particles, %s): loss %.6g vs fp%d oracle %.6g, grad rel err max %.3g, l2 %.3g (%.1fs)"
                    % (k, w["sizes"][k], "+".join(c["tools"]), losses[k], bits, loss_ref, e_max, e_l2,
                       time.perf_counter() - t0))
            grad_checks.append(c)
        worst = max(c["rel_err_max_fp32"] for c in grad_checks)
        gate = dict(bound=GRAD_GATE, vs="fp32 oracle (ti.f32 arithmetic)", rel_err_max=worst,
                    rel_err_l2=max(c["rel_err_l2_fp32"] for c in grad_checks),
                    fp64_rel_err_max=max(c["rel_err_max_fp64"] for c in grad_checks),
                    fp64_rel_err_l2=max(c["rel_err_l2_fp64"] for c in grad_checks),
                    status="pass" if worst <= GRAD_GATE else "FAIL")
        log("gradient gate (<= %.0e vs the fp32 oracle): %s, worst max-element %.3g; fp64 distance %.3g"
            % (GRAD_GATE, gate["status"], worst, gate["fp64_rel_err_max"]))

    # ---- batched path: warmup (includes graph capture), then K timed iterations
    for _ in range(args.warmup):
        opt.iteration()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        opt.iteration()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000 / args.steps
    value = n_total * T / (ms_step / 1000)
    log("batched: %.2f ms/iteration, %.4g particle-substeps/s (fwd+adj)" % (ms_step, value))

    # ---- forward only (restart, 200 substeps, mass grids), its own graph
    opt.forward()
    fwd_ms = timed(opt.forward, max(1, args.steps))
    forward = dict(ms_per_pass=round(fwd_ms, 3), value=n_total * T / (fwd_ms / 1000), unit="particle-substeps/s (forward)")
    log("batched forward only: %.2f ms/pass, %.4g particle-substeps/s" % (fwd_ms, forward["value"]))

    # ---- live per-kernel timing (HIP events on the launch stream), eager pass
    e = batch.engine
    e.set_profiling(True)
    opt._device_pass()
    torch.cuda.synchronize()
    prof = e.profile_read(reset=True)
    e.set_profiling(False)
    kernels = {}
    for ph, (ms, cnt) in prof.items():
        if cnt:
            kernels[ph] = dict(total_ms=round(ms, 4), launches=int(cnt), avg_us=round(1000 * ms / cnt, 3))
    dom = max((p for p in kernels if p in ALGO_BYTES), key=lambda p: kernels[p]["total_ms"])
    avg_s = kernels[dom]["avg_us"] * 1e-6
    algo = ALGO_BYTES[dom] * n_total
    achieved = algo / avg_s / 1e9
    roofline = dict(bound="hbm", kernel=dom, achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(dom),
                    algo_bytes_per_launch=int(algo), avg_launch_us=kernels[dom]["avg_us"])
    # the same kernel against the VALU issue rate (one wave64 instruction per
    # SIMD every 4 cycles): its instruction count per launch is from the
    # committed SQ counter run of this workload
    valu_issue = None
    mv = measured_valu(dom)
    if mv is not None:
        ach = mv[0] / avg_s
        valu_issue = dict(kernel=dom, instr_per_launch=int(mv[0]), achieved=round(ach / 1e9, 2), peak=round(mv[1] / 1e9, 2),
                          unit="G wave-instr/s", frac=round(ach / mv[1], 4))
    log("kernels:", json.dumps(kernels))

    # ---- plain per-scene GPU path, same kernels
    plain = None
    if not args.no_plain:
        del opt
        p = PlainTrajOpt(w["scenes"], H, S, targets, lr=1e-2, init_actions=init, ckpt_k=args.ckpt_k, keep_blocks=args.keep_blocks,
                         deterministic=not args.unordered, resort=resort)
        p.iteration()
        pm = timed(p.iteration, 1)
        plain = dict(ms_per_step=round(pm, 3), value=n_total * T / (pm / 1000), speedup=round(pm / ms_step, 3))
        log("plain per-scene: %.2f ms/iteration, batched speedup %.2fx" % (pm, pm / ms_step))
        del p
    launch_used = batch.engine.get_launch()
    del batch
    torch.cuda.synchronize()
    torch.cuda.empty_cache()

    # ---- the other workloads
    extra = None
    if not args.no_extra:
        extra = {}
        extra["config1_forward"] = config1_forward()
        log("config 1 forward: %.2f ms/pass, %.4g particle-substeps/s" % (extra["config1_forward"]["ms_per_pass"],
                                                                          extra["config1_forward"]["value"]))
        extra["config3_memory"] = config3_memory()
        for k, v in extra["config3_memory"]["runs"].items():
            log("config 3 %s: peak %.3f GB (engine workspace %.3f GB), %.2f s fwd+adj"
                % (k, v["peak_bytes"] / 1e9, v["workspace_bytes"] / 1e9, v["seconds"]))
        torch.cuda.empty_cache()
        extra["config5"] = config5_rate()
        log("config 5: %.1f ms fwd+adj, %.4g particle-substeps/s" % (extra["config5"]["ms_per_pass"], extra["config5"]["value"]))

    # ---- CPU baseline: fp32 oracle on the host cores, bounded sample
    cpu = None
    if not args.no_cpu:
        v, used, spent = cpu_baseline(w, args.cpu_budget, HOST_CORES)
        cpu = dict(value=v, unit="particle-substeps/s (fwd+adjoint)", cores=HOST_CORES, kind="port",
                   sample="scenes %s of the workload (%d particles), all %d substeps fwd+adjoint, %.1f s"
                          % (used, int(sum(w["sizes"][i] for i in used)), T, spent))
        log("cpu baseline: %.4g particle-substeps/s on %d threads" % (v, HOST_CORES))
        k0 = int(np.argsort(w["sizes"])[0])
        v1, used1, spent1 = cpu_baseline(w, 0.0, 1, scenes=[k0])
        cpu["one_core"] = dict(value=v1, cores=1, sample="scene %d (%d particles), all %d substeps fwd+adjoint, %.1f s"
                                                         % (k0, int(w["sizes"][k0]), T, spent1))
        log("cpu baseline, one core: %.4g particle-substeps/s" % v1)

    out = {
        "metric": METRIC, "value": value, "unit": "particle-substeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched trajectory optimisation: %d scenes, particles log-uniform [2k,40k] (total %d), "
                               "64^3 grid each, 1-2 capsule/box tools, %d substeps fwd+adjoint + Adam step"
                               % (len(w["scenes"]), n_total, T),
                   "scenes": len(w["scenes"]), "particles": n_total, "substeps": T, "grid": 64,
                   "ckpt_k": args.ckpt_k or "store-all", "keep_blocks": args.keep_blocks,
                   "launch": launch_used, "autotuned": launch is not None, "hip_graph": not args.no_graph,
                   "fixed_order": not args.unordered, "resort_substeps": resort_used},
        "grad_gate": gate, "grad_checks": grad_checks,
        "plain": plain, "forward": forward, "extra": extra,
        "kernels": kernels, "roofline": roofline, "valu_issue": valu_issue, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
