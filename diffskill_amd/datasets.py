"""DiffSkill's cached-state dataset formats.

A dataset directory (``ENV.cached_state_path`` of a task YAML) holds
- ``init/state_{i}.xz``: an lzma-compressed pickle (protocol 4) of
  ``TaichiEnv.get_state()``, i.e. ``{'state': [x, v, F, C, tool_0, ...],
  'softness': float, 'is_copy': bool}`` with float64 numpy arrays
  (core/diffskill/gen_init_target/state_generator.py:34-45,
  plb/engine/taichi_env.py:159-165);
- ``target/target_{i}.npy``: the particle positions (N, 3) of a target state
  (state_generator.py:52-58);
- ``target/target_imgs.npy``: rendered target images (state_generator.py:64-67).

``MultitaskPlasticineEnv`` reads them with ``pickle.load`` and ``np.load``
(plb/envs/multitask_env.py:22-32,46-61,70-91).  Here state files are read by
an unpickler that admits only numpy arrays, dtypes and scalars, the bytes
encoder of pickle protocols <= 2 and the plain containers pickle builds
without globals: a file naming any other global is refused before anything in
it runs.  ``.npy`` files are read with
``allow_pickle=False``.
"""
import glob
import importlib
import lzma
import os
import pickle
import re

import numpy as np

# (module, name) globals a state file may reference: numpy's array and scalar
# reconstruction under its 1.x (numpy.core) and 2.x (numpy._core) module names,
# and the bytes encoder that pickle protocols <= 2 use for array buffers
_NUMPY_GLOBALS = {
    ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
    ("numpy", "ndarray"), ("numpy", "dtype"),
}
_DATA_GLOBALS = {("_codecs", "encode")}


def _numpy_global(module, name):
    # numpy 2 keeps numpy.core only as a deprecated alias of numpy._core
    if module.startswith("numpy.core"):
        module = "numpy._core" + module[len("numpy.core"):]
    try:
        return getattr(importlib.import_module(module), name)
    except ImportError:
        return getattr(importlib.import_module(module.replace("numpy._core", "numpy.core")), name)


class _StateUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _NUMPY_GLOBALS:
            return _numpy_global(module, name)
        if (module, name) in _DATA_GLOBALS:
            return getattr(importlib.import_module(module), name)
        raise pickle.UnpicklingError("state files may only hold numpy arrays and plain containers; "
                                     "refusing global %s.%s" % (module, name))


def load_state(path):
    """One ``init/state_{i}.xz`` file -> the dict ``TaichiEnv.set_state(**d)`` takes."""
    with lzma.open(path, "rb") as f:
        d = _StateUnpickler(f).load()
    if not isinstance(d, dict) or "state" not in d:
        raise ValueError("%s is not a TaichiEnv state dict" % path)
    return d


def save_state(path, state):
    """Writes ``state`` (a ``TaichiEnv.get_state()`` dict) like
    StateGenerator.save_init: lzma + pickle protocol 4."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with lzma.open(path, "wb") as f:
        pickle.dump(state, f, protocol=4)


def natural_key(s):
    """Sort key ordering embedded integers numerically (natsort's default)."""
    return [int(t) if t.isdigit() else t for t in re.split(r"(\d+)", s)]


class CachedStates:
    """The ``cached_state_path`` layout read by MultitaskPlasticineEnv."""

    def __init__(self, path):
        self.path = path
        self.init_paths = sorted(glob.glob(os.path.join(path, "init", "state_*.xz")), key=natural_key)
        self.target_paths = sorted(glob.glob(os.path.join(path, "target", "target_[0-9]*.npy")), key=natural_key)

    @property
    def num_inits(self):
        return len(self.init_paths)

    @property
    def num_targets(self):
        return len(self.target_paths)

    def init_state(self, i):
        """``init/state_{i}.xz`` (by its number, as multitask_env.py:88-91 opens it)."""
        return load_state(os.path.join(self.path, "init", "state_%d.xz" % int(i)))

    def target_pc(self, i):
        return np.load(os.path.join(self.path, "target", "target_%d.npy" % int(i)), allow_pickle=False)

    def target_pcs(self):
        """All targets in natural order (load_target_pcs, multitask_env.py:22-30)."""
        return np.array([np.load(p, allow_pickle=False) for p in self.target_paths])

    def target_imgs(self):
        return np.load(os.path.join(self.path, "target", "target_imgs.npy"), allow_pickle=False)

    def save_init(self, i, state):
        save_state(os.path.join(self.path, "init", "state_%d.xz" % int(i)), state)

    def save_target(self, i, x):
        os.makedirs(os.path.join(self.path, "target"), exist_ok=True)
        np.save(os.path.join(self.path, "target", "target_%d.npy" % int(i)), np.asarray(x, np.float64))


def target_mass_grid(points, sim_cfg):
    """The mass grid (n, n, n) float32 of a target point cloud, computed on the
    GPU by compute_grid_m_kernel (mpm_simulator.py:456-466): the target of the
    density L1 loss (plb/engine/losses/loss.py:172-175)."""
    import torch
    from .batch import SceneBatch
    b = SceneBatch([dict(sim=sim_cfg, prims=[], x0=np.asarray(points, np.float64))], 2)
    m = b.grid_mass(0)[0].clone()
    torch.cuda.synchronize()
    return m
