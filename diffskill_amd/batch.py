"""Batched scenes: many independent scenes (each with its own grid and tool
primitives) packed into ONE engine context, so every substep phase is one
launch over all scenes, and a batched trajectory optimiser whose whole
forward + adjoint pass is captured once as a HIP graph and replayed.

This replaces the reference's process-per-environment parallelism
(plb/envs/mp_wrapper.py:46-167) for the differentiable path; the per-scene
semantics are those of ``GradModel`` + ``Solver.solve_one_plan``
(plb/engine/function.py:166-190, plb/optimizer/solver.py:97-152).
"""
import numpy as np
import torch

from .config import derive
from .engine.core import Engine, cell_sort_perm, scene_desc
from .engine.mpm_simulator import collision_pairs, collision_rand
from .engine.primitive.primitives import CLASSES


def _prim_objects(prim_cfgs, softness):
    out = []
    for c in prim_cfgs:
        p = CLASSES[c["shape"]](cfg=c)
        p.softness = float(softness)
        out.append(p)
    return out


class SceneBatch:
    """Scenes packed into one context.

    scenes: list of dicts with keys ``sim`` (SIMULATOR cfg), ``prims``
    (PRIMITIVES cfgs), ``x0`` ((n, 3) positions; v = 0, F = I, C = 0 like
    reset_kernel) and optionally ``prim_states`` (one 8-vector per primitive;
    default: each primitive's init state).  ``keep_blocks``: node blocks of
    each substep's grid kept for the adjoint (-1 default capacity, 0 none).
    """

    def __init__(self, scenes, max_frames, ckpt_k=0, softness=666.0, keep_blocks=-1, deterministic=True):
        self.specs = scenes
        self.consts = [derive(s["sim"]) for s in scenes]
        self.prims = [_prim_objects(s["prims"], softness) for s in scenes]
        descs = [scene_desc(c, len(s["x0"]), len(p)) for c, s, p in zip(self.consts, scenes, self.prims)]
        pdescs = [p.desc() for ps in self.prims for p in ps]
        self.engine = Engine(descs, pdescs, max_frames, ckpt_k, keep_blocks)
        e = self.engine
        e.set_deterministic(bool(deterministic))
        self.n_scenes = len(scenes)
        self.prim_offsets = np.concatenate([[0], np.cumsum([len(p) for p in self.prims])]).astype(int)
        self.n_prims = int(self.prim_offsets[-1])
        # tool-tool collision pairs of every scene (each scene draws its own rand_num)
        pairs, rands = [], []
        for k, s in enumerate(scenes):
            pk = collision_pairs([p.cfg for p in self.prims[k]])
            pairs += [(i + int(self.prim_offsets[k]), j + int(self.prim_offsets[k])) for i, j in pk]
            if pk:
                rands.append(collision_rand(len(pk)))
        if pairs:
            e.set_collisions(pairs, np.concatenate(rands, 0))
        frame0 = torch.zeros((24, e.stride), dtype=torch.float32, device="cuda")
        eye = torch.tensor([1., 0, 0, 0, 1, 0, 0, 0, 1], device="cuda")
        ps0 = torch.zeros((max(1, self.n_prims), 8), dtype=torch.float32, device="cuda")
        # particles stored in cell order; perm[k][i] = caller's index of slot i of scene k
        self.perm = []
        for k, s in enumerate(scenes):
            cols = e.scene_cols(k)
            perm = cell_sort_perm(s["x0"], self.consts[k]["n_grid"])
            self.perm.append(perm)
            frame0[0:3, cols] = torch.as_tensor(np.asarray(s["x0"], np.float32)[perm], device="cuda").t()
            frame0[15:24, cols] = eye[:, None]
            states = s.get("prim_states") or [np.asarray(p.init_state + ((0.0,) if p.state_dim == 7 else ()), np.float64)
                                             for p in self.prims[k]]
            for j, st in enumerate(states):
                v = np.zeros(8)
                v[:len(st)] = st
                ps0[self.prim_offsets[k] + j] = torch.as_tensor(v, dtype=torch.float32)
        self.frame0 = frame0
        self.ps0 = ps0
        mask = torch.zeros((max(1, self.n_prims), 7), dtype=torch.float32, device="cuda")
        for k in range(self.n_scenes):
            for j, p in enumerate(self.prims[k]):
                mask[self.prim_offsets[k] + j, :p.action_dim] = 1
        self.action_mask = mask
        self.restart()

    @property
    def total_particles(self):
        return int(sum(self.engine.counts))

    def restart(self):
        """Frame 0 of every scene back to its initial state (set_state at frame 0)."""
        self.engine.set_frame(0, self.frame0)
        if self.n_prims > 0:
            self.engine.prim_state[0, :self.n_prims].copy_(self.ps0[:self.n_prims])

    def grid_mass(self, f, out=None):
        n = self.consts[0]["n_grid"]
        assert all(c["n_grid"] == n for c in self.consts), "grid_mass output needs one grid size"
        if out is None:
            out = torch.zeros((self.n_scenes, n, n, n), dtype=torch.float32, device="cuda")
        self.engine.grid_mass(f, out)
        return out


class BatchedTrajOpt:
    """Per-scene action sequences optimised with Adam on the density-grid L1
    loss (plb/engine/losses/loss.py:172-175) at the final frame.

    One ``iteration()`` = restart -> H env steps x ``substeps`` forward
    substeps (all scenes) -> mass grid -> L1 adjoint -> all adjoint substeps
    -> action adjoints -> Adam step.  With ``use_graph`` the whole iteration,
    the Adam step included (``capturable`` Adam, its state created up front),
    is captured once as a HIP graph and replayed, so no host work sits
    between iterations.  With ``resort`` (the default) the engine re-sorts
    the particles' slots into cell order at every env step boundary
    (mpm_set_resort, when ``substeps`` is at least 8), or every ``resort``
    substeps when an int is given: the losses and action gradients do not
    depend on slot order, and the tile kernels keep reading and writing
    whole cache lines as the dough deforms.
    """

    def __init__(self, batch, H, substeps, targets, lr=1e-2, init_actions=None, use_graph=True, resort=True):
        self.b = batch
        self.H = int(H)
        self.S = int(substeps)
        self.T = self.H * self.S
        r = self.S if resort is True else int(resort or 0)
        batch.engine.set_resort(r if r >= 8 else 0)
        self.targets = targets.contiguous()
        P = max(1, batch.n_prims)
        a = torch.zeros((self.H, P, 7), dtype=torch.float32, device="cuda") if init_actions is None else \
            init_actions.to("cuda", torch.float32).clone()
        self.actions = torch.nn.Parameter(a)
        self.actions.grad = torch.zeros_like(a)
        self.opt = torch.optim.Adam([self.actions], lr=lr, capturable=bool(use_graph))
        if use_graph:
            # the state Adam would create lazily on its first step, created
            # now so that the captured step only updates it in place
            st = self.opt.state[self.actions]
            st["step"] = torch.zeros((), dtype=torch.float32, device="cuda")
            st["exp_avg"] = torch.zeros_like(a)
            st["exp_avg_sq"] = torch.zeros_like(a)
        n = batch.consts[0]["n_grid"]
        self.mass = torch.zeros((batch.n_scenes, n, n, n), dtype=torch.float32, device="cuda")
        self.table = torch.zeros((self.H, P, 7), dtype=torch.float32, device="cuda")
        self.grads = torch.zeros((self.H, P, 7), dtype=torch.float32, device="cuda")
        self.use_graph = use_graph
        self.graph = None

    def _forward_pass(self):
        e = self.b.engine
        self.b.restart()
        torch.clamp(self.actions.detach(), -1, 1, out=self.table)
        self.table.mul_(self.b.action_mask)
        for s in range(self.H):
            e.set_actions(s, self.S, self.table[s])
        e.substep_fwd(0, self.T)
        self.mass.zero_()
        e.grid_mass(self.T, self.mass)

    def _device_pass(self):
        e = self.b.engine
        self._forward_pass()
        e.clear_grads(self.T)
        e.density_loss_bwd(self.T, self.mass, self.targets)
        e.substep_bwd(0, self.T)
        for s in range(self.H):
            e.action_grad_into(s, self.S, self.grads[s])

    def autotune(self, candidates=None, reps=2):
        """Pick the engine's launch shape (particles per work item, workgroup
        counts) by timing this optimiser's own pass; before graph capture."""
        from .tune import autotune
        if self.graph is not None:
            raise RuntimeError("autotune must run before the pass is captured into a graph")
        return autotune(self.b.engine, self._device_pass, candidates, reps)

    def _step(self):
        torch.mul(self.grads, self.b.action_mask, out=self.actions.grad)
        self.opt.step()

    def _capture(self):
        # warm up the device pass on a side stream (the Adam step is not
        # warmed up: that would move the actions), then capture pass + step
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._device_pass()
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._device_pass()
            self._step()
        self.graph = g

    def forward(self):
        """Forward only: restart, the H env steps of every scene and the
        final mass grids (the loss input), no adjoint and no update; with
        ``use_graph`` captured once as a graph of its own and replayed."""
        if not self.use_graph:
            self._forward_pass()
            return
        if getattr(self, "fwd_graph", None) is None:
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self._forward_pass()
            torch.cuda.current_stream().wait_stream(s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._forward_pass()
            self.fwd_graph = g
        self.fwd_graph.replay()

    def iteration(self):
        if self.use_graph:
            if self.graph is None:
                self._capture()
            self.graph.replay()
        else:
            self._device_pass()
            self._step()

    def losses(self):
        return (self.mass - self.targets).abs().sum(dim=(1, 2, 3))

    def action_grads(self):
        return self.grads * self.b.action_mask


class PlainTrajOpt:
    """The plain per-scene GPU path: one single-scene context per scene, each
    driven with its own launches (no packing, no graph) -- the way the
    reference drives one Taichi scene per process."""

    def __init__(self, scenes, H, substeps, targets, lr=1e-2, init_actions=None, ckpt_k=0, softness=666.0, keep_blocks=-1,
                 deterministic=True, resort=True):
        self.parts = []
        P0 = 0
        for k, sc in enumerate(scenes):
            b = SceneBatch([sc], H * substeps + 1, ckpt_k=ckpt_k, softness=softness, keep_blocks=keep_blocks,
                           deterministic=deterministic)
            ia = None
            if init_actions is not None:
                ia = init_actions[:, P0:P0 + max(1, b.n_prims)]
            P0 += b.n_prims
            self.parts.append(BatchedTrajOpt(b, H, substeps, targets[k:k + 1], lr=lr, init_actions=ia, use_graph=False,
                                             resort=resort))

    @property
    def total_particles(self):
        return sum(p.b.total_particles for p in self.parts)

    def iteration(self):
        for p in self.parts:
            p.iteration()

    def losses(self):
        return torch.cat([p.losses() for p in self.parts])
