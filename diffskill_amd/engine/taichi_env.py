"""TaichiEnv: the scene-assembly and state API of plb/engine/taichi_env.py
(initialize, step, get_state / set_state, get_obs) over the HIP engine.

Rendering, the pykeops / geomloss Sinkhorn loss and the Taichi loss module
are not part of this engine.  Initial particles come from the ``SHAPES`` list
(shapes.Shapes, plb/engine/shapes/shape_maker.py) or from ``init_particles``
passed explicitly.
"""
import numpy as np
import torch

from ..config import CfgNode, get_cfg_defaults
from .mpm_simulator import MPMSimulator
from .primitive import Primitives
from .shapes import Shapes


class TaichiEnv:
    def __init__(self, cfg, nn=False, loss=True, return_dist=False, init_particles=None, ckpt_k=None, deterministic=True):
        self.return_dist = bool(return_dist)
        full = get_cfg_defaults()
        full.merge(cfg)
        self.full_cfg = full
        self.cfg = full.ENV
        self.env_name = self.cfg.get("env_name", None)
        self.primitives = Primitives(full.PRIMITIVES)
        if init_particles is None:
            # TaichiEnv.__init__ (taichi_env.py:65-66): Shapes(cfg.SHAPES).get()
            self.shapes = Shapes(full.SHAPES)
            init_particles, self.particle_colors = self.shapes.get()
        else:
            self.shapes = None
            self.particle_colors = None
        self.init_particles = np.asarray(init_particles, np.float64)
        self.n_particles = len(self.init_particles)
        self.simulator = MPMSimulator(full.SIMULATOR, self.primitives, ckpt_k=ckpt_k, deterministic=deterministic)
        self.dim = self.simulator.dim
        self._is_copy = True
        self.device = "cuda"
        self.has_loss = loss

    def set_copy(self, is_copy: bool):
        self._is_copy = is_copy

    def initialize(self, cfg=None, target_path=None):
        """taichi_env.py:103-121: a new cfg rebuilds the particles from its
        SHAPES and updates the primitives' cfgs."""
        if cfg is not None:
            full = get_cfg_defaults()
            full.merge(cfg)
            self.full_cfg = full
            self.cfg = full.ENV
            self.shapes = Shapes(full.SHAPES)
            self.init_particles, self.particle_colors = self.shapes.get()
            self.primitives.update_cfgs(full.PRIMITIVES)
        if self.has_loss and target_path is not None:
            self.load_target_x(target_path)
        self.n_particles = len(self.init_particles)
        self.simulator.initialize(self.n_particles)
        self.primitives.initialize()
        self.simulator.reset(self.init_particles)

    def load_target_x(self, path):
        """taichi_env.py:123-125: the target particle positions of a dataset's
        ``target/target_{i}.npy`` (read without unpickling)."""
        self.target_x = np.load(path, allow_pickle=False)
        self.tensor_target_x = torch.as_tensor(self.target_x, dtype=torch.float32, device=self.device)

    def step(self, action=None):
        if action is not None:
            action = np.array(action)
        self.simulator.step(is_copy=self._is_copy, action=action)

    def get_state(self):
        assert self.simulator.cur == 0
        return {"state": self.simulator.get_state(0), "softness": self.primitives.get_softness(), "is_copy": self._is_copy}

    def set_state(self, state, softness=None, is_copy=None, **kwargs):
        if softness is None:
            softness = state["softness"]
            is_copy = state["is_copy"]
            state = state["state"]
        self.n_particles = len(state[0])
        self.simulator.cur = 0
        self.simulator.set_state(0, state)
        self.primitives.set_softness(softness)
        self._is_copy = is_copy

    def set_primitive_state(self, state, softness, is_copy, **kwargs):
        self.simulator.set_primitive_state(0, state)

    def get_obs(self, s, device):
        """taichi_env.py:356-372 (positions only, plus tool states); with
        return_dist one distance column per primitive (taichi_env.py:85-98,
        374-383): the sum of both jaw distances for a Gripper, the primitive's
        SDF otherwise (for Gripper2 the nearer jaw, Gripper2.sdf)."""
        f = s * self.simulator.substeps
        x = self.simulator.get_torch_x(f, device)
        c = torch.zeros((len(self.primitives), 8), device=device)
        for idx, p in enumerate(self.primitives):
            st = p.get_state(f)
            c[idx, :len(st)] = torch.as_tensor(st, dtype=torch.float32)
        if self.return_dist:
            d = self.simulator.compute_min_dist(f).to(device)
            cols, k = [], 0
            for p in self.primitives:
                if p.shape in ("Gripper", "Gripper2"):
                    pair = d[:, k:k + 2]
                    cols.append(pair.sum(1, keepdim=True) if p.shape == "Gripper" else pair.min(1, keepdim=True).values)
                    k += 2
                else:
                    cols.append(d[:, k:k + 1])
                    k += 1
            x = torch.cat([x] + cols, 1)
        return x.clone(), c.clone()
