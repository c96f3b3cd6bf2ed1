"""Tool primitives: host-side mirror of plb/engine/primitive (primive_base.py,
primitives.py).

A primitive holds its configuration on the host; its per-frame state (pos,
quat wxyz, gap), per-substep velocity and their adjoints live in the engine
context's device buffers once the primitive is bound to an ``MPMSimulator``.
SDF, normal, collision and kinematics run in the HIP kernels
(csrc/mpm_device.h); nothing here computes physics.
"""
import numpy as np
import torch

from ... import _lib
from ...config import CfgNode


def _tuple(v, n=None, fill=0.0):
    v = tuple(float(x) for x in (v if isinstance(v, (list, tuple)) else (v,)))
    if n is not None:
        v = (v + (fill,) * n)[:n]
    return v


class _PrimField:
    """One per-frame field of a primitive (primive_base.py:37-52): ``position``,
    ``rotation``, ``gap`` in the frame state, ``v``, ``w``, ``gap_vel`` in the
    per-substep velocity, read and written through the engine's buffers;
    ``field[f]``, ``field[f] = value``, ``to_numpy()``, ``fill(value)``, and the
    adjoint as ``field.grad`` with the same methods."""

    def __init__(self, prim, which, cols, grad=False):
        self.prim, self.which, self.cols, self.is_grad = prim, which, cols, grad

    @property
    def grad(self):
        return _PrimField(self.prim, self.which, self.cols, True)

    def _buf(self):
        sim = self.prim._sim
        if sim is None or sim.engine is None:
            raise RuntimeError("primitive fields live in the simulator's buffers: initialize the simulator first")
        e = sim.engine
        if self.which == "state":
            return e.prim_state_grad if self.is_grad else e.prim_state
        return e.prim_vel_grad if self.is_grad else e.prim_vel

    def __getitem__(self, f):
        v = self._buf()[int(f), self.prim._index, self.cols[0]:self.cols[1]].detach().double().cpu().numpy()
        return v[0] if len(v) == 1 else v

    def __setitem__(self, f, value):
        b = self._buf()
        b[int(f), self.prim._index, self.cols[0]:self.cols[1]] = torch.as_tensor(
            np.asarray(value, np.float32).reshape(-1), device=b.device)

    def to_numpy(self):
        v = self._buf()[:, self.prim._index, self.cols[0]:self.cols[1]].detach().double().cpu().numpy()
        return v[:, 0] if v.shape[1] == 1 else v

    def fill(self, value):
        self._buf()[:, self.prim._index, self.cols[0]:self.cols[1]] = float(value)


class _ActionBuffer:
    """``action_buffer`` (primive_base.py:49): the clipped action of each env
    step s as ``action_buffer[s]``; its adjoint ``action_buffer.grad[s]``
    (get_action_grad, primive_base.py:276-282)."""

    def __init__(self, prim, grad=False):
        self.prim, self.is_grad = prim, grad

    @property
    def grad(self):
        return _ActionBuffer(self.prim, True)

    def __getitem__(self, s):
        p = self.prim
        if self.is_grad:
            return p.get_action_grad(int(s), 1)[0]
        return np.asarray(p._sim.primitives._table(int(s))[p._index, :p.action_dim], np.float64)

    def fill(self, value):
        if self.is_grad and float(value) == 0.0:
            return  # action adjoints are derived from the velocity adjoints, which the grad fills clear
        raise NotImplementedError("action_buffer.fill: set actions with set_action")


class Primitive:
    """Primitive (primive_base.py:13-329)."""

    state_dim = 7
    shape = None

    def __init__(self, cfg=None, dim=3, max_timesteps=1024, dtype="float32", **kwargs):
        base = self.default_config()
        if cfg is not None:
            base.merge(cfg)
        if kwargs:
            base.merge(kwargs)
        self.cfg = base
        self.dim = dim
        self.max_timesteps = max_timesteps
        self.dtype = dtype
        self.pos_dim = dim
        self.rotation_dim = 4
        self.action_dim = int(self.cfg.action.dim)
        self.friction = float(self.cfg.friction)
        # Primitive.softness is a zero-initialised field until set_softness (primive_base.py:33)
        self.softness = 0.0
        self._sim = None
        self._index = None
        self._host_states = {}
        self._xyz_limit = (_tuple(self.cfg.lower_bound, 3), _tuple(self.cfg.upper_bound, 3))
        # per-frame fields and their adjoints (engine buffers once bound)
        self.position = _PrimField(self, "state", (0, 3))
        self.rotation = _PrimField(self, "state", (3, 7))
        self.v = _PrimField(self, "vel", (0, 3))
        self.w = _PrimField(self, "vel", (3, 6))
        self.action_buffer = _ActionBuffer(self)

    # ------------------------------------------------------------ config
    @classmethod
    def default_config(cls):
        """Primitive.default_config (primive_base.py:314-329)."""
        return CfgNode(dict(shape="", init_pos=(0.3, 0.3, 0.3), init_rot=(1., 0., 0., 0.), color=(0.3, 0.3, 0.3),
                            lower_bound=(0., 0., 0.), upper_bound=(1., 1., 1.), friction=0.9,
                            collision_group=[0., 0., 0.], action=dict(dim=0, scale=())))

    def shape_params(self):
        return {}

    def desc(self):
        """Parameters for the C ABI (mpm_prim_desc)."""
        d = _lib.PrimDesc()
        d.type = _lib.SHAPE_CODES[type(self).shape]
        d.action_dim = self.action_dim
        d.friction = self.friction
        d.softness = self.softness
        p = dict(size=(0.1, 0.1, 0.1), h=0.0, r=0.0, ph=(0.1, 0.1), prot=(1., 0., 0., 0.), min_gap=0.0, max_gap=1.0)
        p.update(self.shape_params())
        for i, v in enumerate(_tuple(p["size"], 3)):
            d.size[i] = v
        d.h, d.r = float(p["h"]), float(p["r"])
        for i, v in enumerate(_tuple(p["ph"], 2)):
            d.ph[i] = v
        for i, v in enumerate(_tuple(p["prot"], 4)):
            d.prot[i] = v
        lo, hi = self._xyz_limit
        for i in range(3):
            d.lo[i], d.hi[i] = lo[i], hi[i]
        d.min_gap, d.max_gap = float(p["min_gap"]), float(p["max_gap"])
        scale = _tuple(self.cfg.action.scale) if self.action_dim > 0 else ()
        for i, v in enumerate(scale[:7]):
            d.action_scale[i] = v
        return d

    @property
    def init_state(self):
        return tuple(self.cfg.init_pos) + tuple(self.cfg.init_rot)

    # ------------------------------------------------------------ state
    def _bind(self, sim, index):
        self._sim = sim
        self._index = index
        for f, s in sorted(self._host_states.items()):
            self._write(f, s)
        self._host_states = {}

    def _write(self, f, s8):
        self._sim._prim_state_view()[f, self._index] = torch.as_tensor(np.asarray(s8, np.float32), device="cuda")

    def _read(self, f):
        return self._sim._prim_state_view()[f, self._index].detach().cpu().numpy().astype(np.float64)

    def get_state(self, f):
        """pos (3) + quat wxyz (4) [+ gap] as float64 (primive_base.py:196-199)."""
        if self._sim is None or self._sim._ctx is None:
            s = self._host_states.get(f, np.zeros(8))
        else:
            s = self._read(f)
        return np.asarray(s[:self.state_dim], np.float64)

    def set_state(self, f, state):
        """primive_base.py:201-204: missing trailing entries keep their value."""
        ss = np.zeros(8)
        ss[:self.state_dim] = self.get_state(f)
        state = np.asarray(state, np.float64).reshape(-1)
        ss[:len(state)] = state
        if self._sim is None or self._sim._ctx is None:
            self._host_states[f] = ss
        else:
            self._write(f, ss)

    def get_state_tensor(self, f):
        return torch.as_tensor(self.get_state(f)[:7], dtype=torch.float32, device="cuda")

    def initialize(self):
        """primive_base.py:229-236."""
        self.set_state(0, self.init_state)
        self.friction = float(self.cfg.friction)
        self._xyz_limit = (_tuple(self.cfg.lower_bound, 3), _tuple(self.cfg.upper_bound, 3))
        if self._sim is not None:
            self._sim._update_prims()

    def update_cfg(self, cfg):
        self.cfg.merge(cfg)
        self.set_state(0, self.init_state)

    def set_position(self, pos):
        s = self.get_state(0)
        s[:3] = pos
        self.set_state(0, s)

    def set_rotation(self, rot):
        s = self.get_state(0)
        s[3:7] = rot
        self.set_state(0, s)

    # ------------------------------------------------------------ actions
    def set_action(self, s, n_substeps, action):
        """Primitive.set_action (primive_base.py:270-274) for this primitive only."""
        if self.action_dim > 0:
            self._sim.primitives._set_one_action(self._index, s, n_substeps, action)

    def get_action_grad(self, s, n):
        """get_action_grad (primive_base.py:276-282): (n, action_dim) float64."""
        if self.action_dim == 0:
            return None
        g = self._sim._action_grad_rows(s, n)
        return g[:, self._index, :self.action_dim].astype(np.float64)


class Sphere(Primitive):
    shape = "Sphere"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.radius = 1.
        return c

    def shape_params(self):
        return dict(r=float(self.cfg.radius))


class Capsule(Primitive):
    shape = "Capsule"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.h = 0.06
        c.r = 0.03
        return c

    def shape_params(self):
        return dict(h=float(self.cfg.h), r=float(self.cfg.r))


class RollingPin(Capsule):
    shape = "RollingPin"


class RollingPinExt(Capsule):
    shape = "RollingPinExt"


class _GapMixin:
    state_dim = 8

    @property
    def gap(self):
        """Per-frame gap (primitives.py:433) and its adjoint."""
        return _PrimField(self, "state", (7, 8))

    @property
    def gap_vel(self):
        return _PrimField(self, "vel", (6, 7))

    @property
    def init_state(self):
        return tuple(self.cfg.init_pos) + tuple(self.cfg.init_rot) + (float(self.cfg.init_gap),)

    def set_state(self, f, state):
        state = np.asarray(state, np.float64).reshape(-1)
        assert len(state) == 8
        Primitive.set_state(self, f, state)


class Chopsticks(_GapMixin, Capsule):
    shape = "Chopsticks"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.h = 0.06
        c.r = 0.03
        c.minimal_gap = 0.06
        c.init_gap = 0.06
        return c

    def shape_params(self):
        return dict(h=float(self.cfg.h), r=float(self.cfg.r), min_gap=float(self.cfg.minimal_gap))


class Box(Primitive):
    shape = "Box"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.size = (0.1, 0.1, 0.1)
        return c

    def shape_params(self):
        return dict(size=_tuple(self.cfg.size, 3))

    def initialize(self, cached_state_path=""):
        super().initialize()


class Gripper(_GapMixin, Box):
    shape = "Gripper"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.size = (0.03, 0.06, 0.03)
        c.minimal_gap = 0.06
        c.maximal_gap = 1.
        c.init_gap = 0.06
        c.round = 0
        return c

    def shape_params(self):
        return dict(size=_tuple(self.cfg.size, 3), min_gap=float(self.cfg.minimal_gap), max_gap=float(self.cfg.maximal_gap))


class Gripper2(_GapMixin, Capsule):
    shape = "Gripper2"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.h = 0.06
        c.r = 0.015
        c.minimal_gap = 0.06
        c.maximal_gap = 1.
        c.init_gap = 0.06
        c.round = 0
        return c

    def shape_params(self):
        return dict(h=float(self.cfg.h), r=float(self.cfg.r), min_gap=float(self.cfg.minimal_gap),
                    max_gap=float(self.cfg.maximal_gap))


class Prism(Primitive):
    shape = "Prism"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.h = (0.1, 0.1)
        c.prot = (1.0, 0.0, 0.0, 0.0)
        return c

    def shape_params(self):
        return dict(ph=_tuple(self.cfg.h, 2), prot=_tuple(self.cfg.prot, 4))


class Knife(Primitive):
    """Knife = max(prism SDF, box SDF), an intersection (primitives.py:769-773)."""
    shape = "Knife"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.h = (0.1, 0.1)
        c.size = (0.1, 0.1, 0.1)
        c.prot = (1.0, 0.0, 0.0, 0.0)
        return c

    def shape_params(self):
        return dict(ph=_tuple(self.cfg.h, 2), prot=_tuple(self.cfg.prot, 4), size=_tuple(self.cfg.size, 3))


class Cylinder(Primitive):
    """Cylinder (primitives.py:293-326): h the disk radius, r the half height."""
    shape = "Cylinder"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.h = 0.2
        c.r = 0.1
        return c

    def shape_params(self):
        return dict(h=float(self.cfg.h), r=float(self.cfg.r))


class Torus(Primitive):
    """Torus (primitives.py:329-356): tx the ring radius, ty the tube radius."""
    shape = "Torus"

    @classmethod
    def default_config(cls):
        c = Primitive.default_config()
        c.tx = 0.2
        c.ty = 0.1
        return c

    def shape_params(self):
        return dict(h=float(self.cfg.tx), r=float(self.cfg.ty))


CLASSES = {c.shape: c for c in (Sphere, Capsule, RollingPin, RollingPinExt, Chopsticks, Box, Gripper, Gripper2,
                                Prism, Knife, Cylinder, Torus)}


class Primitives:
    """Primitives container (primitives.py:820-897)."""

    def __init__(self, cfgs, max_timesteps=1024):
        self.primitives = []
        self.action_dims = [0]
        for c in cfgs:
            c = CfgNode(c) if not isinstance(c, CfgNode) else c
            if c.shape not in CLASSES:
                raise NotImplementedError("primitive shape %r is not one of the reference's classes %s"
                                          % (c.shape, sorted(CLASSES)))
            p = CLASSES[c.shape](cfg=c, max_timesteps=max_timesteps)
            self.primitives.append(p)
            self.action_dims.append(self.action_dims[-1] + p.action_dim)
        self.n = len(self.primitives)
        self._sim = None
        self._step_actions = {}

    def update_cfgs(self, cfgs):
        for p, c in zip(self.primitives, cfgs):
            p.update_cfg(c)

    @property
    def action_dim(self):
        return self.action_dims[-1]

    @property
    def state_dim(self):
        return sum(i.state_dim for i in self.primitives)

    @property
    def state_dims(self):
        return [i.state_dim for i in self.primitives]

    def _table(self, s):
        t = self._step_actions.get(s)
        if t is None:
            t = np.zeros((max(1, self.n), 7), np.float32)
            self._step_actions[s] = t
        return t

    def _upload_actions(self, s, n_substeps):
        t = torch.as_tensor(self._table(s), device="cuda")
        self._sim._set_actions(s, n_substeps, t)

    def _set_one_action(self, j, s, n_substeps, action):
        a = action.detach().float().cpu().numpy() if isinstance(action, torch.Tensor) else np.asarray(action, np.float32)
        a = np.asarray(a, np.float32).reshape(-1)
        self._table(s)[j, :] = 0
        self._table(s)[j, :len(a)] = a
        self._upload_actions(s, n_substeps)

    def set_action(self, s, n_substeps, action):
        """primitives.py:863-867: clip to [-1, 1], one slice per primitive."""
        if isinstance(action, torch.Tensor):
            action = action.detach().double().cpu().numpy()
        action = np.asarray(action).reshape(-1).clip(-1, 1)
        assert len(action) == self.action_dims[-1]
        t = self._table(s)
        t[:] = 0
        for i in range(self.n):
            a = action[self.action_dims[i]:self.action_dims[i + 1]]
            t[i, :len(a)] = a
        if self.n > 0:
            self._upload_actions(s, n_substeps)

    def get_grad(self, n):
        grads = [g for g in (p.get_action_grad(0, n) for p in self.primitives) if g is not None]
        return np.concatenate(grads, axis=1)

    def set_softness(self, softness=666.):
        for p in self.primitives:
            p.softness = float(softness)
        if self._sim is not None:
            self._sim._update_prims()

    def get_softness(self):
        return self.primitives[0].softness

    def __getitem__(self, item):
        if isinstance(item, tuple):
            item = item[0]
        return self.primitives[item]

    def __len__(self):
        return len(self.primitives)

    def __iter__(self):
        return iter(self.primitives)

    def initialize(self, cached_state_path=""):
        for p in self.primitives:
            p.initialize()
