"""Seeded synthetic scenes shared by the tests: one spec builds both the
product objects (reference-style cfg dicts) and the oracle's plain dicts.

The oracle parameters are written out here from the primitive cfg
independently of the product's own descriptor code."""
import numpy as np

from diffskill_amd.config import derive

from oracle import prim_from_cfg, prim_state_from_cfg, sim_from_cfg


def oracle_prim(cfg, softness=666.0):
    return prim_from_cfg(cfg, softness)


def oracle_sim(sim_cfg):
    """(oracle constants, product-derived constants); the two derivations are
    independent and must agree."""
    o, substeps = sim_from_cfg(sim_cfg)
    c = derive(sim_cfg)
    assert c["substeps"] == substeps and c["n_grid"] == o["n_grid"] and abs(c["dt"] - o["dt"]) < 1e-15
    return o, c


def prim_state0(cfg):
    return prim_state_from_cfg(cfg)


def particles_box(n, lo, hi, seed=0, v_sigma=0.0, F_jitter=0.0):
    rng = np.random.default_rng(seed)
    st = np.zeros((n, 24))
    st[:, :3] = rng.uniform(lo, hi, (n, 3))
    if v_sigma:
        st[:, 3:6] = rng.normal(0, v_sigma, (n, 3))
    F = np.tile(np.eye(3).reshape(1, 9), (n, 1))
    if F_jitter:
        F = F + rng.normal(0, F_jitter, (n, 9))
    st[:, 15:24] = F
    return st


def state_list(st, prim_states):
    n = st.shape[0]
    return [st[:, 0:3].copy(), st[:, 3:6].copy(), st[:, 15:24].reshape(n, 3, 3).copy(),
            st[:, 6:15].reshape(n, 3, 3).copy()] + [np.asarray(p, np.float64) for p in prim_states]


def capsule_scene(n=2000, seed=0, n_grid=32, dt=2e-4, substeps=5, max_steps=64):
    sim = dict(yield_stress=50.0, E=5e3, nu=0.2, ground_friction=1.5, gravity=(0, -9.8, 0), n_grid=n_grid, dt=dt,
               substeps=substeps, max_steps=max_steps, n_particles=n)
    prim = dict(shape="Capsule", h=0.1, r=0.03, init_pos=(0.5, 0.17, 0.5), init_rot=(0.926321, 0.201374, 0.302061, 0.100687),
                friction=0.9, action=dict(dim=3, scale=(0.01, 0.01, 0.01)))
    st = particles_box(n, [0.4, 0.05, 0.4], [0.6, 0.15, 0.6], seed=seed, v_sigma=0.1)
    return sim, [prim], st
