"""Engine context: one C-ABI context over a torch-owned device workspace.

The workspace is a single uint8 CUDA (HIP) tensor of ``mpm_workspace_bytes``;
every buffer the context carves out of it is exposed here as a zero-copy
torch view (frames, adjoints, primitive frames).  All work is enqueued on
torch's current stream.
"""
import ctypes

import numpy as np
import torch

from .. import _lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def cell_sort_perm(x, n_grid):
    """Permutation that orders particles by (4x4x4 cell block, cell), the order
    the engine bins them in each substep (cell index base = int(x / dx - 0.5),
    truncating, like the reference's stencil base).  Kernels read particles
    through per-substep index lists; storing them in this order keeps those
    reads local."""
    x = np.asarray(x, np.float64)
    b = np.clip(np.trunc(x * n_grid - 0.5).astype(np.int64), 0, n_grid - 1)
    nb = (n_grid + 3) // 4
    blk = ((b[:, 0] >> 2) * nb + (b[:, 1] >> 2)) * nb + (b[:, 2] >> 2)
    cell = ((b[:, 0] & 3) << 4) | ((b[:, 1] & 3) << 2) | (b[:, 2] & 3)
    return np.argsort(blk * 64 + cell, kind="stable")


def scene_desc(c, n_particles, n_prims):
    """mpm_scene_desc from the constants of ``config.derive``."""
    d = _lib.SceneDesc()
    d.n_grid = int(c["n_grid"])
    d.n_particles = int(n_particles)
    d.n_prims = int(n_prims)
    for k in ("dx", "inv_dx", "dt", "p_vol", "p_mass", "mu", "lam", "yield_stress", "ground_friction", "lower_bound"):
        setattr(d, k, float(c[k]))
    for i in range(3):
        d.gravity[i] = float(c["gravity"][i])
    return d


class Engine:
    """A packed multi-scene context (scene k: constants, particle count, primitives)."""

    def __init__(self, scene_descs, prim_descs, max_frames, ckpt_k=0, keep_blocks=-1):
        if not torch.cuda.is_available():
            raise RuntimeError("the MI355X engine needs a HIP device (torch.cuda.is_available() is False)")
        self.L = _lib.load()
        self.n_scenes = len(scene_descs)
        self.max_frames = int(max_frames)
        self.ckpt_k = int(ckpt_k or 0)
        # node blocks of each substep's grid kept for the adjoint (-1 default, 0 none)
        self.keep_blocks = int(keep_blocks)
        self._scenes = (_lib.SceneDesc * self.n_scenes)(*scene_descs)
        self.n_prims = len(prim_descs)
        self._prims = (_lib.PrimDesc * max(1, self.n_prims))(*prim_descs)
        nbytes = self.L.mpm_workspace_bytes(self._scenes, self.n_scenes, self.max_frames, self.ckpt_k, self.keep_blocks)
        if nbytes <= 0:
            raise RuntimeError("mpm_workspace_bytes failed")
        self.ws = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        ctx = ctypes.c_void_p()
        _lib.check(self.L.mpm_create(self._scenes, self.n_scenes, self._prims, self.max_frames, self.ckpt_k,
                                     self.keep_blocks, ctypes.c_void_p(self.ws.data_ptr()), int(nbytes), _stream(), ctypes.byref(ctx)),
                   "mpm_create")
        self.ctx = ctx
        self.stride = self.L.mpm_stride(ctx)
        self.offsets = [self.L.mpm_particle_offset(ctx, k) for k in range(self.n_scenes)]
        self.counts = [int(d.n_particles) for d in scene_descs]
        self.n_grids = [int(d.n_grid) for d in scene_descs]
        ps, psg, pv, pvg = (ctypes.c_void_p() for _ in range(4))
        _lib.check(self.L.mpm_prim_buffers(ctx, ctypes.byref(ps), ctypes.byref(psg), ctypes.byref(pv), ctypes.byref(pvg)),
                   "mpm_prim_buffers")
        np_ = max(1, self.n_prims)
        shape = (self.max_frames, np_, 8)
        self.prim_state = self._view(ps.value, shape)
        self.prim_state_grad = self._view(psg.value, shape)
        self.prim_vel = self._view(pv.value, shape)
        self.prim_vel_grad = self._view(pvg.value, shape)

    def __del__(self):
        ctx = getattr(self, "ctx", None)
        if ctx is not None and ctx.value:
            try:
                self.L.mpm_destroy(ctx)
            except Exception:
                pass
            self.ctx = None

    # ---------------------------------------------------------------- views
    def _view(self, ptr, shape):
        n = 1
        for s in shape:
            n *= s
        off = int(ptr) - self.ws.data_ptr()
        assert 0 <= off and off + 4 * n <= self.ws.numel(), "buffer outside the workspace"
        return self.ws[off:off + 4 * n].view(torch.float32).view(*shape)

    def frame(self, f):
        """(24, stride) view of frame f (recomputed from its checkpoint if needed)."""
        p = ctypes.c_void_p()
        _lib.check(self.L.mpm_frame(self.ctx, int(f), _stream(), ctypes.byref(p)), "mpm_frame")
        return self._view(p.value, (24, self.stride))

    def adjoint(self, f):
        p = ctypes.c_void_p()
        _lib.check(self.L.mpm_adjoint(self.ctx, int(f), ctypes.byref(p)), "mpm_adjoint")
        return self._view(p.value, (24, self.stride))

    def scene_cols(self, k):
        o = self.offsets[k]
        return slice(o, o + self.counts[k])

    # ---------------------------------------------------------------- calls
    def set_frame(self, f, soa):
        soa = soa.contiguous()
        assert soa.shape == (24, self.stride) and soa.dtype == torch.float32 and soa.is_cuda
        _lib.check(self.L.mpm_set_frame(self.ctx, int(f), ctypes.c_void_p(soa.data_ptr()), _stream()), "mpm_set_frame")

    def reset_frame0(self, x_soa):
        x_soa = x_soa.contiguous()
        assert x_soa.shape == (3, self.stride)
        _lib.check(self.L.mpm_reset_frame0(self.ctx, ctypes.c_void_p(x_soa.data_ptr()), _stream()), "mpm_reset_frame0")

    def copy_frame(self, src, dst):
        _lib.check(self.L.mpm_copy_frame(self.ctx, int(src), int(dst), _stream()), "mpm_copy_frame")

    def set_actions(self, s, n_substeps, table):
        table = table.contiguous().float()
        assert table.is_cuda and table.shape[-1] == 7
        _lib.check(self.L.mpm_set_actions(self.ctx, int(s), int(n_substeps), ctypes.c_void_p(table.data_ptr()), _stream()),
                   "mpm_set_actions")

    def substep_fwd(self, s, count=1):
        _lib.check(self.L.mpm_substep_fwd(self.ctx, int(s), int(count), _stream()), "mpm_substep_fwd")

    def substep_bwd(self, s, count=1):
        _lib.check(self.L.mpm_substep_bwd(self.ctx, int(s), int(count), _stream()), "mpm_substep_bwd")

    def _mass_grid(self, t, name, out=False):
        """t as a contiguous fp32 device grid of every scene's n^3 nodes
        (scene after scene), checked before the kernel indexes it."""
        n = sum(g ** 3 for g in self.n_grids)
        if t.numel() != n:
            raise ValueError("%s: %d values, the scenes' mass grids hold %d" % (name, t.numel(), n))
        if out:
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise ValueError("%s: needs a contiguous float32 device tensor" % name)
            return t
        return t.to(device="cuda", dtype=torch.float32).contiguous()

    def grid_mass(self, f, out):
        out = self._mass_grid(out, "grid_mass", out=True)
        _lib.check(self.L.mpm_grid_mass(self.ctx, int(f), ctypes.c_void_p(out.data_ptr()), _stream()), "mpm_grid_mass")

    def grid_mass_bwd(self, f, g):
        g = self._mass_grid(g, "grid_mass_bwd")
        _lib.check(self.L.mpm_grid_mass_bwd(self.ctx, int(f), ctypes.c_void_p(g.data_ptr()), _stream()), "mpm_grid_mass_bwd")

    def density_loss_bwd(self, f, mass, target, weight=1.0):
        """Adjoint of weight * sum |mass - target| (the reference's density
        loss, losses/loss.py:173-175, :213) into frame f's particles, the
        node adjoint weight * sign(mass - target) formed in the grid-mass
        adjoint kernel."""
        mass = self._mass_grid(mass, "density_loss_bwd mass")
        target = self._mass_grid(target, "density_loss_bwd target")
        _lib.check(self.L.mpm_density_loss_bwd(self.ctx, int(f), ctypes.c_void_p(mass.data_ptr()), ctypes.c_void_p(target.data_ptr()),
                                               float(weight), _stream()), "mpm_density_loss_bwd")

    def add_obs_grad(self, f, xv_slots=None, prim_grad=None):
        a = ctypes.c_void_p(xv_slots.data_ptr()) if xv_slots is not None else None
        b = ctypes.c_void_p(prim_grad.data_ptr()) if prim_grad is not None else None
        _lib.check(self.L.mpm_add_obs_grad(self.ctx, int(f), a, b, _stream()), "mpm_add_obs_grad")

    def min_dist_cols(self, scene=-1):
        n = self.L.mpm_min_dist_cols(self.ctx, int(scene))
        if n < 0:
            _lib.check(n, "mpm_min_dist_cols")
        return n

    def min_dist(self, f):
        """compute_min_dist of frame f: (n_cols, stride) float32, slot order."""
        nc = max(1, self.min_dist_cols())
        out = torch.zeros((nc, self.stride), dtype=torch.float32, device="cuda")
        _lib.check(self.L.mpm_min_dist(self.ctx, int(f), ctypes.c_void_p(out.data_ptr()), nc, _stream()), "mpm_min_dist")
        return out

    def min_dist_bwd(self, f, g_slots):
        """Adds the adjoint of the distances (n_cols, stride) into frame f's adjoints."""
        g = g_slots.detach().to("cuda", torch.float32).contiguous()
        _lib.check(self.L.mpm_min_dist_bwd(self.ctx, int(f), ctypes.c_void_p(g.data_ptr()), int(g.shape[0]), _stream()),
                   "mpm_min_dist_bwd")

    def set_collisions(self, pairs, rand_num):
        """Tool-tool collision pairs [(i, j), ...] (primitive indices of this
        context, grouped by scene) and their rand_num (n_pairs, n_points, 3);
        an empty list switches the projection off."""
        pairs = np.ascontiguousarray(np.asarray(pairs, np.int32).reshape(-1, 2))
        n = pairs.shape[0]
        if n == 0:
            _lib.check(self.L.mpm_set_collisions(self.ctx, 0, None, 1, None, None, 0, _stream()), "mpm_set_collisions")
            self.coll_buf = None
            self.n_pairs = 0
            return
        rand = np.ascontiguousarray(np.asarray(rand_num, np.float32))
        if rand.ndim != 3 or rand.shape[0] != n or rand.shape[2] != 3:
            raise ValueError("rand_num must be (n_pairs, n_points, 3), got %s" % (rand.shape,))
        nbytes = self.L.mpm_collision_bytes(self.ctx, n, rand.shape[1])
        if nbytes <= 0:
            raise ValueError("mpm_collision_bytes rejected %d pairs x %d points" % (n, rand.shape[1]))
        buf = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        _lib.check(self.L.mpm_set_collisions(self.ctx, n, pairs.ctypes.data_as(ctypes.c_void_p), rand.shape[1],
                                             rand.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(buf.data_ptr()), int(nbytes),
                                             _stream()), "mpm_set_collisions")
        self.coll_buf = buf        # the context reads it until the next call
        self.n_pairs = n

    def get_collisions(self, f):
        """Records of frame f (substep f-1 -> f): (collision_idx (n_pairs,),
        chosen points (n_pairs, 3), positions of i before each move (n_pairs, 3))."""
        n = getattr(self, "n_pairs", 0)
        idx = np.full(n, -1, np.int32)
        pts = np.zeros((n, 3), np.float32)
        pre = np.zeros((n, 3), np.float32)
        if n:
            _lib.check(self.L.mpm_get_collisions(self.ctx, int(f), idx.ctypes.data_as(ctypes.c_void_p),
                                                 pts.ctypes.data_as(ctypes.c_void_p), pre.ctypes.data_as(ctypes.c_void_p),
                                                 _stream()), "mpm_get_collisions")
        return idx, pts, pre

    def action_grad(self, s, n_substeps):
        out = torch.zeros((max(1, self.n_prims), 7), dtype=torch.float32, device="cuda")
        _lib.check(self.L.mpm_action_grad(self.ctx, int(s), int(n_substeps), ctypes.c_void_p(out.data_ptr()), _stream()),
                   "mpm_action_grad")
        return out

    def action_grad_into(self, s, n_substeps, out):
        """Action adjoint of env step s written into ``out`` ([n_prims, 7] float32, contiguous)."""
        assert out.is_cuda and out.dtype == torch.float32 and out.is_contiguous() and out.shape[-1] == 7
        _lib.check(self.L.mpm_action_grad(self.ctx, int(s), int(n_substeps), ctypes.c_void_p(out.data_ptr()), _stream()),
                   "mpm_action_grad")

    def clear_grads(self, f=-1):
        _lib.check(self.L.mpm_clear_grads(self.ctx, int(f), _stream()), "mpm_clear_grads")

    def set_launch(self, particles_per_item=0, tile_ctas=0, grid_ctas=0):
        """Launch shape of the particle and grid phases (0 keeps a value)."""
        _lib.check(self.L.mpm_set_launch(self.ctx, int(particles_per_item), int(tile_ctas), int(grid_ctas)), "mpm_set_launch")

    def get_launch(self):
        v = [ctypes.c_int() for _ in range(3)]
        _lib.check(self.L.mpm_get_launch(self.ctx, *[ctypes.byref(x) for x in v]), "mpm_get_launch")
        return dict(particles_per_item=v[0].value, tile_ctas=v[1].value, grid_ctas=v[2].value)

    def set_deterministic(self, on=True):
        """Fixed-order execution: bitwise-identical results for identical
        inputs and launch shape (see mpm_set_deterministic)."""
        _lib.check(self.L.mpm_set_deterministic(self.ctx, 1 if on else 0), "mpm_set_deterministic")

    @property
    def deterministic(self):
        return self.L.mpm_get_deterministic(self.ctx) == 1

    def set_resort(self, substeps):
        """Re-sort the particles' slots into cell order every ``substeps``
        substeps (0: off); invalidates every frame after 0 (see
        mpm_set_resort).  The engine's buffer for it is held here."""
        substeps = int(substeps)
        buf = None
        if substeps:
            nbytes = self.L.mpm_resort_bytes(self.ctx, substeps)
            _lib.check(0 if nbytes > 0 else -1, "mpm_resort_bytes")
            buf = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        _lib.check(self.L.mpm_set_resort(self.ctx, substeps, ctypes.c_void_p(buf.data_ptr() if buf is not None else 0),
                                         int(buf.numel() if buf is not None else 0)), "mpm_set_resort")
        self._resort_buf = buf

    @property
    def resort(self):
        return self.L.mpm_get_resort(self.ctx)

    def slot_order(self, f):
        """int64 tensor (stride,): the slot in frame f of the particle at each
        slot of frame 0."""
        out = torch.empty(self.stride, dtype=torch.int32, device="cuda")
        _lib.check(self.L.mpm_slot_order(self.ctx, f, out.data_ptr(), _stream()), "mpm_slot_order")
        return out.long()

    def set_profiling(self, on=True):
        _lib.check(self.L.mpm_set_profiling(self.ctx, 1 if on else 0), "mpm_set_profiling")

    def profile_read(self, reset=True):
        """{phase: (total_ms, launches)} from the HIP events recorded around each kernel."""
        n = len(_lib.PHASES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_longlong * n)()
        _lib.check(self.L.mpm_profile_read(self.ctx, ms, cnt, 1 if reset else 0), "mpm_profile_read")
        return {p: (ms[i], cnt[i]) for i, p in enumerate(_lib.PHASES)}

    def update_prims(self, prim_descs):
        self._prims = (_lib.PrimDesc * max(1, self.n_prims))(*prim_descs)
        _lib.check(self.L.mpm_update_prims(self.ctx, self._prims, _stream()), "mpm_update_prims")
