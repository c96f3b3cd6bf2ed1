"""Dataset formats through the GPU engine: a TaichiEnv state written to a
DiffSkill state file and read back into another environment, a target point
cloud loaded as TaichiEnv.target_x, and its mass grid against the oracle."""
import numpy as np
import pytest

from oracle import Oracle, reset_state
from scenes import oracle_sim, particles_box, prim_state0, state_list

pytestmark = pytest.mark.gpu

SIM = dict(yield_stress=50.0, E=5e3, nu=0.2, ground_friction=1.5, gravity=(0, -9.8, 0), n_grid=32, dt=2e-4,
           substeps=3, max_steps=16, n_particles=1500)
PRIMS = [dict(shape="Gripper", size=(0.015, 0.04, 0.03), init_pos=(0.5, 0.2, 0.5), init_rot=(1.0, 0.0, 0.0, 0.0),
              init_gap=0.12, minimal_gap=0.02, friction=1.0, action=dict(dim=7, scale=(0.01,) * 7)),
         dict(shape="Box", size=(0.04, 0.02, 0.04), init_pos=(0.5, 0.3, 0.5), init_rot=(1.0, 0.0, 0.0, 0.0),
              friction=5.0, action=dict(dim=6, scale=(0.01,) * 6))]


def _env(st):
    from diffskill_amd.engine import TaichiEnv
    env = TaichiEnv(dict(SIMULATOR=SIM, PRIMITIVES=PRIMS, SHAPES=[]), init_particles=st[:, :3])
    env.initialize()
    return env


def test_state_file_round_trip_through_taichi_env(tmp_path):
    from diffskill_amd import datasets
    print("case dataset state round trip", flush=True)
    st = particles_box(1500, [0.4, 0.05, 0.4], [0.6, 0.12, 0.6], seed=3, v_sigma=0.1, F_jitter=0.02)
    env = _env(st)
    env.set_state(state_list(st, [prim_state0(p) for p in PRIMS]), softness=666.0, is_copy=True)
    env.step(np.random.default_rng(0).uniform(-1, 1, 13))      # copy mode: the result lands in frame 0
    s0 = env.get_state()
    cs = datasets.CachedStates(str(tmp_path / "ds"))
    cs.save_init(0, s0)
    env2 = _env(particles_box(1500, [0.3, 0.05, 0.3], [0.5, 0.1, 0.5], seed=9))
    env2.set_state(**datasets.CachedStates(str(tmp_path / "ds")).init_state(0))
    s1 = env2.get_state()
    assert s1["softness"] == s0["softness"] and s1["is_copy"] == s0["is_copy"]
    for a, b in zip(s0["state"], s1["state"]):
        np.testing.assert_array_equal(a, b)


def test_target_point_cloud_and_its_mass_grid(tmp_path):
    from diffskill_amd import datasets
    print("case dataset target grid", flush=True)
    pts = particles_box(4000, [0.35, 0.05, 0.45], [0.65, 0.09, 0.55], seed=4)[:, :3]
    cs = datasets.CachedStates(str(tmp_path / "ds"))
    cs.save_target(0, pts)
    env = _env(particles_box(1500, [0.4, 0.05, 0.4], [0.6, 0.12, 0.6], seed=3))
    env.initialize(target_path=str(tmp_path / "ds" / "target" / "target_0.npy"))
    np.testing.assert_array_equal(env.target_x, pts)
    assert tuple(env.tensor_target_x.shape) == (4000, 3)
    m = datasets.target_mass_grid(pts, SIM).cpu().double().numpy()
    osim, _ = oracle_sim(SIM)
    ref = Oracle(osim, []).grid_mass(reset_state(pts))
    assert abs(m.sum() - 4000 * osim["p_mass"]) < 1e-5 * 4000 * osim["p_mass"]
    np.testing.assert_allclose(m, ref, rtol=0, atol=1e-5 * ref.max())
