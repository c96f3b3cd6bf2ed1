"""Seeded synthetic workloads for the five benchmark configurations of the
differentiable plasticine simulator (material E = 5e3, nu = 0.2, von Mises
yield 50, density 1, gravity -9.8 y, 64^3 grid unless stated).

Seeds: particles 0, actions 1, targets 2, scene sizes 3, pile placement 4.
Scenes sit on the floor just above the lower wall band (3 nodes).
"""
import numpy as np

DENSITY = 30000 / 0.1 ** 3   # particles per unit volume (plb/engine/shapes/shape_maker.py:56-57)


def material(**kw):
    d = dict(dim=3, quality=1, quality_multiplier=1., E=5e3, nu=0.2, yield_stress=50., ground_friction=1.5,
             gravity=(0., -9.8, 0.), lower_bound=0., dtype="float32")
    d.update(kw)
    return d


def _floor(n_grid):
    return 3.0 / n_grid + 0.002


def _slab_target(n, footprint_area, center_xz, y0, rng, thickness=0.02):
    """n points uniform in a slab of the given thickness and square footprint."""
    L = np.sqrt(footprint_area)
    p = rng.random((n, 3))
    p[:, 0] = center_xz[0] + (p[:, 0] - 0.5) * L
    p[:, 2] = center_xz[1] + (p[:, 2] - 0.5) * L
    p[:, 1] = y0 + p[:, 1] * thickness
    return np.clip(p, 3.0 / 64, 1 - 3.0 / 64)


def config1(n=5000, n_env_steps=50, substeps=10, dt=5e-4):
    """Plain forward: one 5k-particle dough block U([.4,.6]x[.05,.15]x[.4,.6]),
    capsule r 0.03 (segment h = 0.2), friction 0.9, N(0, 0.01^2) actions per env step."""
    rng = np.random.default_rng(0)
    x0 = rng.uniform([0.4, 0.05, 0.4], [0.6, 0.15, 0.6], (n, 3))
    sim = material(n_particles=n, dt=dt, substeps=substeps, max_steps=n_env_steps * substeps + 1)
    prim = dict(shape="Capsule", h=0.2, r=0.03, init_pos=(0.5, 0.19, 0.5), init_rot=(0.7071068, 0., 0., 0.7071068),
                friction=0.9, action=dict(dim=3, scale=(0.01, 0.01, 0.01)))
    acts = np.random.default_rng(1).normal(0, 0.01, (n_env_steps, 3))
    return dict(sim=sim, prims=[prim], x0=x0, actions=acts, H=n_env_steps, substeps=substeps)


def config2(n=20000, n_env_steps=25, substeps=20):
    """Single differentiable scene: 20k particles in an ellipsoid (0.1, 0.05, 0.1),
    RollingPin (rot-y, rot-long-axis, translate-y) actions U(-0.02, 0.02),
    500 substeps at dt 1e-4, loss L1(mass grid, 0.02-thick slab of the same mass)."""
    rng = np.random.default_rng(0)
    fl = _floor(64)
    c = np.array([0.5, fl + 0.05, 0.5])
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= rng.random((n, 1)) ** (1 / 3)
    x0 = c + p * np.array([0.1, 0.05, 0.1])
    sim = material(n_particles=n, substeps=substeps, max_steps=n_env_steps * substeps + 1)
    prim = dict(shape="RollingPin", h=0.3, r=0.03, init_pos=(0.5, fl + 0.1 + 0.033, 0.5),
                init_rot=(0.7071068, 0.7071068, 0., 0.), friction=0.9,
                action=dict(dim=3, scale=(0.6666666666666667, 0.06666666666666668, 0.01)))
    acts = np.random.default_rng(1).uniform(-0.02, 0.02, (n_env_steps, 3))
    vol = 4 / 3 * np.pi * 0.1 * 0.05 * 0.1
    tgt = _slab_target(n, vol / 0.02, (0.5, 0.5), fl, np.random.default_rng(2))
    return dict(sim=sim, prims=[prim], x0=x0, actions=acts, H=n_env_steps, substeps=substeps, target_x=tgt)


def config4(n_scenes=64, n_env_steps=10, substeps=20, lo=2000, hi=40000):
    """Batched trajectory optimisation: 64 scenes, particle counts log-uniform
    in [2k, 40k] (seed 3), each a dough cube at the reference density with its
    own 64^3 grid and 1-2 tools (a horizontal capsule roller, optionally a box
    pusher); 200 substeps (10 env steps x 20) per iteration; loss: L1 between
    the mass grid and a 0.02-thick slab of the same mass."""
    rs = np.random.default_rng(3)
    sizes = np.round(np.exp(rs.uniform(np.log(lo), np.log(hi), n_scenes))).astype(int)
    fl = _floor(64)
    scenes, targets, acts = [], [], []
    for k, n in enumerate(sizes):
        r = np.random.default_rng(1000 + k)
        a = (n / DENSITY) ** (1 / 3)
        cx, cz = 0.5 + r.uniform(-0.08, 0.08), 0.5 + r.uniform(-0.08, 0.08)
        x0 = np.array([cx, fl + a / 2, cz]) + (r.random((n, 3)) - 0.5) * a
        prims = [dict(shape="Capsule", h=0.2, r=0.03, init_pos=(cx, fl + a + 0.033, cz),
                      init_rot=(0.7071068, 0., 0., 0.7071068), friction=0.9,
                      action=dict(dim=3, scale=(0.01, 0.01, 0.01)))]
        if r.random() < 0.5:
            prims.append(dict(shape="Box", size=(0.015, 0.03, 0.03), init_pos=(cx + a / 2 + 0.03, fl + 0.032, cz),
                              friction=0.9, action=dict(dim=3, scale=(0.01, 0.01, 0.01))))
        sim = material(n_particles=int(n), substeps=substeps, max_steps=n_env_steps * substeps + 1)
        scenes.append(dict(sim=sim, prims=prims, x0=x0))
        targets.append(_slab_target(int(n), a ** 3 / 0.02, (cx, cz), fl, np.random.default_rng(2000 + k)))
        acts.append(np.random.default_rng(100 + k).uniform(-0.3, 0.3, (n_env_steps, len(prims), 3)))
    return dict(scenes=scenes, target_x=targets, actions=acts, H=n_env_steps, substeps=substeps, sizes=sizes)


def action_tensor(acts_per_scene, H):
    """Stack per-scene (H, n_prims_k, 3) actions into the packed (H, P, 7) table."""
    P = sum(a.shape[1] for a in acts_per_scene)
    out = np.zeros((H, max(1, P), 7), np.float32)
    o = 0
    for a in acts_per_scene:
        out[:, o:o + a.shape[1], :a.shape[2]] = a
        o += a.shape[1]
    return out
