"""TaichiEnv: the scene-assembly and state API of plb/engine/taichi_env.py
(initialize, step, get_state / set_state, get_obs) over the HIP engine.

Rendering, pykeops / geomloss point-cloud losses and the Taichi loss module
are not part of this engine.  Initial particles come from ``SHAPES`` boxes and
spheres sampled like plb/engine/shapes/shape_maker.py, or from
``init_particles`` passed explicitly.
"""
import numpy as np
import torch

from ..config import CfgNode, get_cfg_defaults
from .mpm_simulator import MPMSimulator
from .primitive import Primitives


def sample_shapes(shapes, seed=None):
    """Box / sphere sampling of shape_maker.py:81-105 (30000 particles per 0.1^3)."""
    rng = np.random.RandomState(seed) if seed is not None else np.random
    objs = []
    for s in shapes:
        s = CfgNode(s) if not isinstance(s, CfgNode) else s
        kind = s.get("shape", "box")
        if kind == "box":
            width = np.array(s.get("width", (0.1, 0.1, 0.1)), dtype=np.float64)
            n = s.get("n_particles", None) or max(int(np.prod(width) / (0.1 ** 3) * 30000), 1)
            p = (rng.random_sample((int(n), 3)) * 2 - 1) * (0.5 * width) + np.array(s.init_pos)
        elif kind == "sphere":
            r = float(s.get("radius", 0.05))
            n = s.get("n_particles", None) or max(int((r ** 3) * 4 * np.pi / 3 / (0.1 ** 3) * 30000), 1)
            p = rng.normal(size=(int(n), 3))
            p /= np.linalg.norm(p, axis=-1, keepdims=True)
            u = rng.random_sample(size=(int(n), 1)) ** (1. / 3)
            p = p * u * r + np.array(s.init_pos)
        else:
            raise NotImplementedError("shape %r is not supported by this sampler" % kind)
        objs.append(p)
    return np.concatenate(objs).astype(np.float64)


class TaichiEnv:
    def __init__(self, cfg, nn=False, loss=True, return_dist=False, init_particles=None, ckpt_k=None):
        self.return_dist = bool(return_dist)
        full = get_cfg_defaults()
        full.merge(cfg)
        self.full_cfg = full
        self.cfg = full.ENV
        self.env_name = self.cfg.get("env_name", None)
        self.primitives = Primitives(full.PRIMITIVES)
        if init_particles is None:
            init_particles = sample_shapes(full.SHAPES)
        self.init_particles = np.asarray(init_particles, np.float64)
        self.n_particles = len(self.init_particles)
        self.simulator = MPMSimulator(full.SIMULATOR, self.primitives, ckpt_k=ckpt_k)
        self.dim = self.simulator.dim
        self._is_copy = True
        self.device = "cuda"
        self.has_loss = loss

    def set_copy(self, is_copy: bool):
        self._is_copy = is_copy

    def initialize(self, cfg=None, target_path=None):
        """taichi_env.py:103-121."""
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
