"""Drop-in engine API mirroring plb/engine (MPMSimulator, Primitives, GradModel, TaichiEnv)."""
from .mpm_simulator import MPMSimulator
from .primitive import Primitives
from .function import GradModel
from .taichi_env import TaichiEnv

__all__ = ["MPMSimulator", "Primitives", "GradModel", "TaichiEnv"]
