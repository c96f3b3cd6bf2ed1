"""MPMSimulator: drop-in for plb/engine/mpm_simulator.py backed by the HIP engine.

Same constructor, attributes and methods as the reference class (substep,
substep_grad, step, get_state/set_state, reset, get_x/get_v/get_torch_x,
compute_grid_m_kernel and its gradient).  The per-frame particle fields of
the reference (``x, v, C, F [max_steps, N]``) are frames of the engine
context: every frame is kept (store-all) unless ``ckpt_k`` asks for
checkpoints every K frames with recomputation in the backward pass.
"""
import numpy as np
import torch

from ..config import derive
from .core import Engine, cell_sort_perm, scene_desc


NUM_PARTICLE_COLLISION = 600  # points sampled per collision pair (mpm_simulator.py:59)


def collision_pairs(prim_cfgs):
    """set_num_collision (mpm_simulator.py:69-80): (i, j) for every primitive
    i and j with collision_group[j] > 0 in i's cfg, i-major.  Entries past
    the end of a collision_group list (the default has three) count as 0."""
    pairs = []
    n = len(prim_cfgs)
    for i in range(n):
        grp = prim_cfgs[i].get("collision_group", None) if hasattr(prim_cfgs[i], "get") else None
        if grp is None:
            continue
        for j in range(n):
            if j < len(grp) and float(grp[j]) > 0:
                pairs.append((i, j))
    return pairs


def collision_rand(n_pairs, n_points=NUM_PARTICLE_COLLISION):
    """rand_num of MPMSimulator.initialize (mpm_simulator.py:87-96): the
    legacy generator seeded with 42, uniform(-1.5, 1.5, (pairs, points, 3)),
    stored as float32; the caller's global numpy random state is untouched."""
    return np.random.RandomState(42).uniform(-1.5, 1.5, (n_pairs, n_points, 3)).astype(np.float32)


class _Scalar:
    """Stand-in for a 0-d Taichi field read as ``field[None]``."""

    def __init__(self, v=0):
        self.v = v

    def __getitem__(self, k):
        return self.v

    def __setitem__(self, k, v):
        self.v = v


class _ParticleField:
    """A per-frame particle field of MPMSimulator (x, v, F or C, ``[frame][particle]``,
    mpm_simulator.py:40-43), read from the engine's frames in the caller's
    particle order: ``field[f, p]``, ``field.to_numpy()`` (frames not simulated
    read as zeros).  ``field.grad.fill(0)`` clears the engine's adjoints, what
    GradModel.reset does for every field (function.py:44-61); ``field.grad[f, p]``
    reads the adjoint of the frame the backward pass is at."""

    _INDEX = {"x": 0, "v": 1, "F": 2, "C": 3}

    def __init__(self, sim, name, grad=False):
        self.sim, self.name, self.is_grad = sim, name, grad

    @property
    def grad(self):
        return _ParticleField(self.sim, self.name, True)

    def _frame(self, f):
        k = self._INDEX[self.name]
        if self.is_grad:
            x, v, F, C = self.sim.get_adjoint(f)
            return (x, v, F, C)[k]
        return self.sim._frame_tensors(f)[k]

    def __getitem__(self, key):
        f, p = key
        return self._frame(int(f))[int(p)].detach().double().cpu().numpy()

    def to_numpy(self):
        n = self.sim.n_particles[None]
        shape = (3,) if self.name in ("x", "v") else (3, 3)
        out = np.zeros((self.sim.max_steps, n) + shape)
        for f in range(self.sim.max_steps):
            try:
                out[f] = self._frame(f).detach().double().cpu().numpy()
            except RuntimeError:
                pass
        return out

    def fill(self, value):
        if self.is_grad and float(value) == 0.0:
            self.sim.clear_grads()
            return
        raise NotImplementedError("%s.fill: set particle states with set_state / reset" % self.name)


class _GridField:
    """``grid_m`` with the field methods callers use (fill / to_torch / to_numpy / grad.from_torch)."""

    def __init__(self, sim):
        self.sim = sim
        self.data = None
        self.grad = _GridGrad(self)

    def _ensure(self):
        n = self.sim.n_grid
        if self.data is None:
            self.data = torch.zeros((n, n, n), dtype=torch.float32, device="cuda")
        return self.data

    def fill(self, v):
        self._ensure().fill_(float(v))

    def to_torch(self, device=None):
        t = self._ensure().clone()
        return t.to(device) if device is not None else t

    def to_numpy(self):
        return self._ensure().detach().cpu().numpy()


class _GridGrad:
    def __init__(self, field):
        self.field = field
        self.data = None

    def from_torch(self, t):
        self.data = t.detach().to("cuda", torch.float32).contiguous()

    def fill(self, v):
        n = self.field.sim.n_grid
        self.data = torch.full((n, n, n), float(v), dtype=torch.float32, device="cuda")


class MPMSimulator:
    def __init__(self, cfg, primitives=(), ckpt_k=None, keep_blocks=-1, deterministic=True):
        c = derive(cfg)
        self.cfg = cfg
        self.consts = c
        self.dim = c["dim"]
        self.dtype = c["dtype"]
        self._yield_stress = c["yield_stress"]
        self.ground_friction = c["ground_friction"]
        self.default_gravity = c["gravity"]
        self.lower_bound = c["lower_bound"]
        self.n_primitive = len(primitives)
        self.n_grid = c["n_grid"]
        self.dx, self.inv_dx = c["dx"], c["inv_dx"]
        self.dt = c["dt"]
        self.p_vol, self.p_rho = c["p_vol"], c["p_rho"]
        self.p_mass = c["p_mass"]
        self._mu, self._lam = c["mu"], c["lam"]
        self.max_steps = c["max_steps"]
        self.substeps = c["substeps"]
        self.res = (self.n_grid,) * 3
        self.primitives = primitives
        self.n_particles = _Scalar(0)
        self.cur = 0
        if ckpt_k is None:
            ckpt_k = int(cfg.get("ckpt_k", 0) or 0) if hasattr(cfg, "get") else 0
        self.ckpt_k = ckpt_k
        self.keep_blocks = keep_blocks
        # fixed-order execution of the engine (bitwise-reproducible results;
        # False: unordered float atomics)
        self.deterministic = bool(deterministic)
        self.grid_m = _GridField(self)
        self.x = _ParticleField(self, "x")
        self.v = _ParticleField(self, "v")
        self.F = _ParticleField(self, "F")
        self.C = _ParticleField(self, "C")
        # particles are stored in cell order: slot i holds the caller's particle perm[i]
        self._perm = None
        self._inv = None
        self.engine = None
        self._ctx = None
        self.num_particle_collision = NUM_PARTICLE_COLLISION
        self.set_num_collision()
        self.rand_num = collision_rand(self.num_collision) if self.num_collision > 0 else None
        if hasattr(primitives, "_sim"):
            primitives._sim = self
        for j, p in enumerate(primitives):
            p._sim = self
            p._index = j

    # ---------------------------------------------------------------- collisions
    def set_num_collision(self):
        """mpm_simulator.py:69-80."""
        self.collision_pairs = [list(p) for p in collision_pairs([p.cfg for p in self.primitives])]
        self.num_collision = len(self.collision_pairs)
        return self.num_collision

    def get_collisions(self, f):
        """collision_idx of frame f (substep f-1 -> f) per pair (-1: no
        contact), the chosen surface points (n_pairs, 3) and the positions of
        the projected primitive before each pair's move."""
        return self.engine.get_collisions(f)

    # ---------------------------------------------------------------- context
    def _prim_descs(self):
        return [p.desc() for p in self.primitives]

    def _ensure_ctx(self, n):
        n = int(n)
        if self.engine is not None and self.engine.counts[0] == n:
            return
        keep = None
        if self.engine is not None and self.n_primitive > 0:
            keep = self.engine.prim_state[0, :self.n_primitive].clone()
        self.engine = Engine([scene_desc(self.consts, n, self.n_primitive)], self._prim_descs(), self.max_steps, self.ckpt_k,
                             self.keep_blocks)
        self._ctx = self.engine.ctx
        self.engine.set_deterministic(self.deterministic)
        if keep is not None:
            self.engine.prim_state[0, :self.n_primitive] = keep
        for j, p in enumerate(self.primitives):
            p._bind(self, j)
        if self.num_collision > 0:
            self.engine.set_collisions(self.collision_pairs, self.rand_num)

    def _prim_state_view(self):
        return self.engine.prim_state

    def _update_prims(self):
        if self.engine is not None:
            self.engine.update_prims(self._prim_descs())

    def _set_actions(self, s, n_substeps, table):
        self.engine.set_actions(s, n_substeps, table)

    def _action_grad_rows(self, s, n):
        rows = [self.engine.action_grad(s + i, self.substeps).cpu().numpy() for i in range(n)]
        return np.stack(rows, 0)

    def _cols(self):
        return self.engine.scene_cols(0)

    def _set_perm(self, x):
        perm = cell_sort_perm(x, self.n_grid)
        inv = np.empty_like(perm)
        inv[perm] = np.arange(len(perm))
        self._perm = torch.as_tensor(perm, device="cuda")
        self._inv = torch.as_tensor(inv, device="cuda")
        return perm

    def _to_caller(self, t):
        """rows in slot order -> rows in the caller's particle order"""
        return t if self._inv is None else t.index_select(0, self._inv)

    def _to_slots(self, t):
        return t if self._perm is None else t.index_select(0, self._perm)

    # ---------------------------------------------------------------- setup
    def initialize(self, n_particles):
        """mpm_simulator.py:82-97 (material constants are per scene)."""
        self.n_particles[None] = int(n_particles)
        self._ensure_ctx(n_particles)

    def reset(self, x):
        """reset_kernel: frame 0 = positions, v = 0, F = I, C = 0 (mpm_simulator.py:402-413)."""
        x = np.asarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x, np.float64)
        n = x.shape[0]
        self.n_particles[None] = n
        self._ensure_ctx(n)
        perm = self._set_perm(x)
        soa = torch.zeros((3, self.engine.stride), dtype=torch.float32, device="cuda")
        soa[:, self._cols()] = torch.as_tensor(x[perm], dtype=torch.float32, device="cuda").t()
        self.engine.reset_frame0(soa)
        self.cur = 0

    # ---------------------------------------------------------------- stepping
    def substep(self, s):
        self.engine.substep_fwd(s, 1)

    def substep_grad(self, s):
        self.engine.substep_bwd(s, 1)

    def step(self, is_copy, action=None):
        """mpm_simulator.py:440-451."""
        start = 0 if is_copy else self.cur
        self.cur = start + self.substeps
        if action is not None:
            self.primitives.set_action(start // self.substeps, self.substeps, action)
        self.engine.substep_fwd(start, self.substeps)
        if is_copy:
            self.engine.copy_frame(self.cur, 0)
            self.cur = 0

    # ---------------------------------------------------------------- state io
    def _frame_tensors(self, f):
        fr = self._to_caller(self.engine.frame(f)[:, self._cols()].t())
        n = fr.shape[0]
        x = fr[:, 0:3]
        v = fr[:, 3:6]
        C = fr[:, 6:15].reshape(n, 3, 3)
        F = fr[:, 15:24].reshape(n, 3, 3)
        return x, v, F, C

    def get_state(self, f):
        """[x, v, F, C, prim states...] as float64 numpy (mpm_simulator.py:380-390)."""
        out = [t.detach().double().cpu().numpy().copy() for t in self._frame_tensors(f)]
        for p in self.primitives:
            out.append(p.get_state(f))
        return out

    def set_state(self, f, state):
        """mpm_simulator.py:392-396; n_particles becomes len(state[0])."""
        n = len(state[0])
        self.n_particles[None] = n
        self._ensure_ctx(n)
        soa = torch.zeros((24, self.engine.stride), dtype=torch.float32, device="cuda")
        cols = self._cols()
        perm = self._set_perm(np.asarray(state[0], np.float64).reshape(n, 3))
        x, v, F, C = [torch.as_tensor(np.asarray(a, np.float32)[perm]).to("cuda") for a in state[:4]]
        soa[0:3, cols] = x.reshape(n, 3).t()
        soa[3:6, cols] = v.reshape(n, 3).t()
        soa[6:15, cols] = C.reshape(n, 9).t()
        soa[15:24, cols] = F.reshape(n, 9).t()
        self.engine.set_frame(f, soa)
        for s, p in zip(state[4:], self.primitives):
            p.set_state(f, s)

    def set_primitive_state(self, f, state):
        for s, p in zip(state[4:], self.primitives):
            p.set_state(f, s)

    def get_x(self, f):
        return self._frame_tensors(f)[0].double().cpu().numpy()

    def get_v(self, f):
        return self._frame_tensors(f)[1].double().cpu().numpy()

    def get_torch_x(self, f, device):
        return self._frame_tensors(f)[0].clone().to(device)

    def get_torch_xv(self, f):
        """(N, 6) float32 device tensor [x | v] of frame f (GradModel._get_obs layout)."""
        fr = self.engine.frame(f)[:, self._cols()]
        return self._to_caller(fr[0:6].t()).contiguous()

    # ---------------------------------------------------------------- mass grid
    def compute_grid_m_kernel(self, f):
        """Adds the particle mass grid of frame f into grid_m (mpm_simulator.py:456-466)."""
        g = self.grid_m._ensure()
        self.engine.grid_mass(f, g)

    def clear_and_compute_grid_m(self, f):
        self.grid_m.fill(0)
        self.compute_grid_m_kernel(f)

    def get_m(self):
        return self.grid_m.to_numpy()

    def compute_grid_m_grad(self, f, grad=None):
        """compute_grid_m_kernel.grad(f): adds d(grad . m)/dx into the adjoint of frame f."""
        g = grad if grad is not None else self.grid_m.grad.data
        self.engine.grid_mass_bwd(f, g)

    # ---------------------------------------------------------------- adjoints
    def clear_grads(self):
        """Zero every adjoint (GradModel.reset, function.py:44-61)."""
        if self.engine is not None:
            self.engine.clear_grads(-1)

    def add_obs_grad(self, f, xv_grad=None, prim_grad=None):
        """_set_obs_grad (function.py:130-142): xv_grad (N, 6), prim_grad (n_prim, 8)."""
        xs = None
        if xv_grad is not None:
            xs = torch.zeros((self.engine.stride, 6), dtype=torch.float32, device="cuda")
            xs[self._cols()] = self._to_slots(xv_grad.detach().to("cuda", torch.float32).reshape(-1, 6))
        pg = None
        if prim_grad is not None and self.n_primitive > 0:
            pg = prim_grad.detach().to("cuda", torch.float32).reshape(self.n_primitive, 8).contiguous()
        self.engine.add_obs_grad(f, xs, pg)

    def compute_min_dist(self, f):
        """compute_min_dist (function.py:79-88): (N, n_cols) distances of the
        particles of frame f to every primitive (both jaws of a Gripper or
        Gripper2), caller order."""
        d = self.engine.min_dist(f)[:, self._cols()]
        return self._to_caller(d.t()).contiguous()

    def compute_min_dist_grad(self, f, grad):
        """compute_min_dist.grad (function.py:154-160): grad (N, n_cols) adds into
        the adjoints of frame f (particle x and primitive states)."""
        g = self._to_slots(grad.detach().to("cuda", torch.float32).reshape(self.n_particles[None], -1))
        gs = torch.zeros((g.shape[1], self.engine.stride), dtype=torch.float32, device="cuda")
        gs[:, self._cols()] = g.t()
        self.engine.min_dist_bwd(f, gs)

    def get_adjoint(self, f):
        """(x, v, F, C) adjoints of frame f as (N,3), (N,3), (N,3,3), (N,3,3) tensors."""
        a = self._to_caller(self.engine.adjoint(f)[:, self._cols()].t())
        n = a.shape[0]
        return a[:, 0:3], a[:, 3:6], a[:, 15:24].reshape(n, 3, 3), a[:, 6:15].reshape(n, 3, 3)

    def scale_adjoint(self, f, alpha):
        """decay_kernel (function.py:66-77): scale the adjoints of frame f."""
        self.engine.adjoint(f).mul_(alpha)
        if self.n_primitive > 0:
            self.engine.prim_state_grad[f, :self.n_primitive, :7].mul_(alpha)
