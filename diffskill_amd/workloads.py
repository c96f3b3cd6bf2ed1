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


def config3(n=30000, n_env_steps=50, substeps=20, n_targets=30000):
    """Two-primitive gather scene: n particles in 4 spherical piles (radius
    U[0.03, 0.06], placement seed 4) on the table; a parallel Gripper of two
    box jaws with full extents 0.02 x 0.06 x 0.06 (open/close DoF) and a thin
    spatula Box of full extents 0.12 x 0.005 x 0.12; 1000 substeps at dt 1e-4;
    loss: one-way Chamfer L2 from the particles to a seeded target cloud
    (a single merged pile)."""
    rng = np.random.default_rng(4)
    fl = _floor(64)
    radii = rng.uniform(0.03, 0.06, 4)
    centres = []
    while len(centres) < 4:   # piles apart from each other, inside the work area
        c = rng.uniform(0.3, 0.7, 2)
        r = radii[len(centres)]
        if all(np.hypot(*(c - q)) > r + radii[i] + 0.01 for i, q in enumerate(centres)):
            centres.append(c)
    vol = radii ** 3
    counts = np.floor(n * vol / vol.sum()).astype(int)
    counts[0] += n - counts.sum()
    prng = np.random.default_rng(0)
    pts = []
    for c, r, m in zip(centres, radii, counts):
        d = prng.normal(size=(m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d *= r * prng.random((m, 1)) ** (1 / 3)
        pts.append(np.array([c[0], fl + r, c[1]]) + d)
    x0 = np.concatenate(pts)
    sim = material(n_particles=n, substeps=substeps, max_steps=n_env_steps * substeps + 1)
    cx, cz = np.mean(centres, axis=0)
    gripper = dict(shape="Gripper", size=(0.01, 0.03, 0.03), init_pos=(cx, fl + 0.03, cz), init_rot=(1., 0., 0., 0.),
                   init_gap=0.3, minimal_gap=0.02, maximal_gap=0.5, friction=1.0,
                   action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0., 0., 0., 0.02)))
    spatula = dict(shape="Box", size=(0.06, 0.0025, 0.06), init_pos=(cx, fl + 0.13, cz), friction=5.0,
                   action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0., 0., 0.)))
    arng = np.random.default_rng(1)
    acts = [arng.uniform(-0.3, 0.3, (n_env_steps, 7)), arng.uniform(-0.3, 0.3, (n_env_steps, 6))]
    acts[0][:, 6] = np.abs(acts[0][:, 6])    # the gripper closes
    trng = np.random.default_rng(2)
    rt = (vol.sum()) ** (1 / 3)
    d = trng.normal(size=(n_targets, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= rt * trng.random((n_targets, 1)) ** (1 / 3)
    target = np.array([cx, fl + rt, cz]) + d
    return dict(sim=sim, prims=[gripper, spatula], x0=x0, actions=acts, H=n_env_steps, substeps=substeps,
                target_points=target)


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


def config5(n=1000000, n_env_steps=40, substeps=50, dt=1e-4, ckpt_k=50):
    """Large cutting scene: n particles in a box on a 128^3 grid (dx = 1/128);
    1M particles fill 0.3 x 0.1 x 0.3, smaller n shrink the box at the same
    density (~53 particles per cell).  A Knife (box body of full extents
    0.2 x 0.1 x 0.002 intersected with a triangular-prism blade, edge down)
    starts just above the block and translates down at 0.5 units/s with
    seeded lateral jitter; 2000 substeps fwd + adjoint with a checkpoint every
    50 frames; loss: mass-grid L1 to the block split into two halves offset
    +-0.05 x the block width in x.  dt defaults to 1e-4: at 2.5e-4 (and down
    to 1.25e-4) the explicit elastic step is beyond its stability limit on
    this grid and the oracle diverges."""
    fl = _floor(128)
    rng = np.random.default_rng(0)
    size = np.array([0.3, 0.1, 0.3]) * (n / 1e6) ** (1 / 3)
    lo = np.array([0.5 - size[0] / 2, fl, 0.5 - size[2] / 2])
    x0 = lo + rng.random((n, 3)) * size
    sim = material(n_particles=n, n_grid=128, dt=dt, substeps=substeps, max_steps=n_env_steps * substeps + 1)
    blade_h = 0.04       # prism size: apex 0.04 below the knife centre, edge along x
    knife = dict(shape="Knife", size=(0.1, 0.05, 0.001), h=(blade_h, 0.1), prot=(0., 0.7071068, 0., 0.7071068),
                 init_pos=(0.5, fl + size[1] + blade_h + 0.002, 0.5), friction=0.0,
                 lower_bound=(0., fl + blade_h, 0.), upper_bound=(1., 1., 1.),
                 action=dict(dim=3, scale=(dt * substeps,) * 3))
    arng = np.random.default_rng(1)
    acts = np.zeros((n_env_steps, 3))
    acts[:, 1] = -0.5                                 # 0.5 units/s downwards
    acts[:, [0, 2]] = arng.normal(0, 0.1, (n_env_steps, 2))
    tgt = x0.copy()
    tgt[:, 0] += np.where(x0[:, 0] < 0.5, -1.0, 1.0) * size[0] / 6
    return dict(sim=sim, prims=[knife], x0=x0, actions=acts, H=n_env_steps, substeps=substeps, target_x=tgt,
                ckpt_k=ckpt_k)


def action_tensor(acts_per_scene, H):
    """Stack per-scene (H, n_prims_k, 3) actions into the packed (H, P, 7) table."""
    P = sum(a.shape[1] for a in acts_per_scene)
    out = np.zeros((H, max(1, P), 7), np.float32)
    o = 0
    for a in acts_per_scene:
        out[:, o:o + a.shape[1], :a.shape[2]] = a
        o += a.shape[1]
    return out
