"""diffskill_amd: MI355X-native differentiable MLS-MPM (the DiffSkill plasticine
simulator hot path) -- hand-written HIP kernels for gfx950 behind the
reference's MPMSimulator / Primitives / GradModel / TaichiEnv API."""
from . import config
from ._lib import load as load_library

__all__ = ["config", "load_library"]
