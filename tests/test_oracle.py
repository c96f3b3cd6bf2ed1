"""Pins of the CPU oracle (oracle/mpm_oracle.c) by checks of a different
nature than the restatement itself: the reference's own data files, library
routines (LAPACK SVD, scipy rotations, torch autograd), closed forms,
conservation laws, and central finite differences of the oracle's forward."""
import json
import os

import numpy as np
import pytest

from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg

HERE = os.path.dirname(os.path.abspath(__file__))


def base_sim(**kw):
    d = dict(E=5e3, nu=0.2, yield_stress=50.0, gravity=(0, -9.8, 0), ground_friction=1.5)
    d.update(kw)
    return d


# ---------------------------------------------------------------- constants
def test_derived_constants_match_reference_arithmetic():
    """MPMSimulator.__init__ (mpm_simulator.py:19-39): quality 1 -> 64^3, dt 1e-4,
    substeps int(2e-3 // 1e-4) = 19 (float floor division, not 20), p_vol squared."""
    o, S = sim_from_cfg({})
    assert o["n_grid"] == 64 and o["dt"] == 1e-4 and S == 19
    assert o["p_vol"] == 6.103515625e-5 == o["p_mass"]
    o, S = sim_from_cfg(dict(quality_multiplier=1.25))   # plb/cut/cut_rearrange.yml
    assert o["n_grid"] == 80 and abs(o["dt"] - 8e-5) < 1e-20 and S == 24
    o, S = sim_from_cfg(dict(quality=2))
    assert o["n_grid"] == 128 and o["dt"] == 5e-5 and S == 39
    o, _ = sim_from_cfg(dict(E=5e3, nu=0.2))
    assert abs(o["mu"] - 5e3 / 2.4) < 1e-9 and abs(o["lam"] - 5e3 * 0.2 / (1.2 * 0.6)) < 1e-9


# ---------------------------------------------------------------- reference data files
def test_asset_grids_pin_particle_mass_and_partition_of_unity():
    """Every target grid shipped with the reference (plb/envs/assets, 64^3 float64)
    holds exactly 10000 particles' mass: sum = 10000 * (0.5 dx)^2 -- the squared
    p_vol convention -- and the oracle's mass grid sums to N * p_mass."""
    fx = json.load(open(os.path.join(HERE, "golden", "asset_grids.json")))
    assert len(fx) == 50
    o, _ = sim_from_cfg({})
    for name, d in fx.items():
        assert d["shape"] == [64, 64, 64]
        assert abs(d["sum"] - 10000 * o["p_mass"]) < 1e-12, name
    orc = Oracle(o, [])
    x = np.random.default_rng(0).uniform(0.1, 0.9, (10000, 3))
    m = orc.grid_mass(reset_state(x))
    assert abs(m.sum() - 0.6103515625) < 1e-12


def test_asset_grids_pin_truncating_cell_index():
    """Negative node mass in the reference's grids occurs only at y index 1: the
    signature of base = int(x * inv_dx - 0.5) truncating toward zero for
    particles resting at y < 0.5 dx (mpm_simulator.py:459).  With floor every
    quadratic B-spline weight would be >= 0.  The oracle reproduces it."""
    fx = json.load(open(os.path.join(HERE, "golden", "asset_grids.json")))
    seen = 0
    for name in ("Rope3D-v1.npy", "Rope3D-v5.npy", "Writer3D-v1.npy", "Writer3D-v3.npy", "Writer3D-v4.npy"):
        neg = np.array(fx[name]["negative_nodes"])
        assert len(neg) > 0 and np.all(neg[:, 1] == 1), name
        seen += len(neg)
    assert seen > 0
    o, _ = sim_from_cfg({})
    orc = Oracle(o, [])
    x = np.array([[0.5, 0.0, 0.5], [0.3, 0.003, 0.7]])
    m = orc.grid_mass(reset_state(x))
    neg = np.argwhere(m < 0)
    assert len(neg) > 0 and np.all(neg[:, 1] == 1)
    # at y = 0 exactly: weights (1.125, -0.25, 0.125) along y
    m0 = orc.grid_mass(reset_state(x[:1]))
    wy = m0.sum(axis=(0, 2)) / o["p_mass"]
    np.testing.assert_allclose(wy[:3], [1.125, -0.25, 0.125], atol=1e-12)


# ---------------------------------------------------------------- SVD
def _rot(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


@pytest.mark.parametrize("case", ["random", "inverted", "near_identity", "degenerate", "rank1"])
def test_svd_against_lapack(case):
    rng = np.random.default_rng(hash(case) % 2 ** 32)
    orc = Oracle(sim_from_cfg({})[0], [])
    for _ in range(50):
        if case == "random":
            A = rng.normal(size=(3, 3))
        elif case == "inverted":
            A = _rot(rng) @ np.diag(rng.uniform(0.2, 2, 3) * [1, 1, -1]) @ _rot(rng).T
        elif case == "near_identity":
            A = np.eye(3) + 1e-4 * rng.normal(size=(3, 3))
        elif case == "degenerate":
            A = _rot(rng) @ np.diag([1.3, 1.3, 0.4]) @ _rot(rng).T
        else:
            A = np.outer(rng.normal(size=3), rng.normal(size=3))
        U, S, V = orc.svd(A)
        np.testing.assert_allclose(U @ S @ V.T, A, atol=1e-12)
        np.testing.assert_allclose(U.T @ U, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(V.T @ V, np.eye(3), atol=1e-12)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        s = np.diag(S)
        np.testing.assert_allclose(np.abs(s), np.linalg.svd(A, compute_uv=False), atol=1e-12)
        assert s[0] >= s[1] >= abs(s[2]) - 1e-15
        if abs(np.linalg.det(A)) > 1e-12:
            assert np.sign(s[2]) == np.sign(np.linalg.det(A))


def test_backward_svd_matches_autograd_of_library_svd():
    """backward_svd (mpm_simulator.py:136-156) equals reverse-mode autodiff of a
    library SVD (torch.linalg.svd, fp64) for separated singular values."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(3)
    orc = Oracle(sim_from_cfg({})[0], [])
    for _ in range(20):
        A = _rot(rng) @ np.diag([1.9, 1.2, 0.6]) @ _rot(rng).T
        U, S, V = orc.svd(A)
        gU, gV = rng.normal(size=(3, 3)), rng.normal(size=(3, 3))
        gS = np.diag(rng.normal(size=3))
        ref = orc.backward_svd(gU, gS, gV, U, S, V)
        At = torch.tensor(A, dtype=torch.float64, requires_grad=True)
        Ut, St, Vht = torch.linalg.svd(At)
        # align torch's convention with the oracle's (column signs); loss is invariant in that choice
        sgn = torch.tensor(np.sign(np.sum(U * Ut.detach().numpy(), axis=0)))
        Ut = Ut * sgn
        Vt = Vht.transpose(0, 1) * sgn
        loss = (Ut * torch.tensor(gU)).sum() + (St * torch.tensor(np.diag(gS))).sum() + (Vt * torch.tensor(gV)).sum()
        loss.backward()
        np.testing.assert_allclose(ref, At.grad.numpy(), atol=1e-9)


# ---------------------------------------------------------------- closed forms
def test_von_mises_return_map_closed_form():
    """compute_von_mises (mpm_simulator.py:165-182) on diagonal F: textbook
    radial return in log strain."""
    sim, _ = sim_from_cfg(base_sim())
    orc = Oracle(sim, [])
    sig = np.array([1.4, 0.8, 0.9])
    st = reset_state(np.array([[0.5, 0.5, 0.5]]))
    st[0, 15:24] = np.diag(sig).reshape(-1)
    Fn, _ = orc.particle_stress(st)
    eps = np.log(sig)
    eh = eps - eps.mean()
    n = np.sqrt(eh @ eh + 1e-8)
    dg = n - sim["yield_stress"] / (2 * sim["mu"])
    assert dg > 0
    expect = np.exp(eps - dg / n * eh)
    np.testing.assert_allclose(np.sort(np.diag(Fn.reshape(3, 3))), np.sort(expect), rtol=1e-12)
    np.testing.assert_allclose(Fn.reshape(3, 3) - np.diag(np.diag(Fn.reshape(3, 3))), 0, atol=1e-12)
    # elastic (no yield): F unchanged
    st[0, 15:24] = np.diag([1.001, 0.999, 1.0]).reshape(-1)
    Fn, _ = orc.particle_stress(st)
    np.testing.assert_allclose(Fn.reshape(3, 3), np.diag([1.001, 0.999, 1.0]), atol=1e-15)


def test_p2g_conserves_mass_and_momentum():
    """p2g (mpm_simulator.py:198-225): sum_i m_i = N p_mass; sum_i p_i = p_mass
    sum_p v_p, because the affine term sums to zero (sum_i w_i (x_i - x_p) = 0)."""
    sim, _ = sim_from_cfg(base_sim(n_grid=32, gravity=(0, 0, 0)))
    orc = Oracle(sim, [])
    rng = np.random.default_rng(1)
    st = reset_state(rng.uniform(0.3, 0.7, (500, 3)))
    st[:, 3:6] = rng.normal(size=(500, 3))
    st[:, 6:15] = rng.normal(size=(500, 9))
    st[:, 15:24] += 0.05 * rng.normal(size=(500, 9))
    orc.substep(st, np.zeros((1, 8)), np.zeros((1, 7)))
    g = orc.grid.reshape(-1, 4)
    assert abs(g[:, 3].sum() - 500 * sim["p_mass"]) < 1e-15
    np.testing.assert_allclose(g[:, :3].sum(0), sim["p_mass"] * st[:, 3:6].sum(0), atol=1e-14)


def test_rigid_translation_is_preserved():
    """Uniform velocity, C = 0, F = I, no gravity, no tools, away from walls:
    one substep keeps v, C = 0, F = I and moves x by dt v (APIC exactness for
    constant fields)."""
    sim, _ = sim_from_cfg(base_sim(n_grid=32, gravity=(0, 0, 0)))
    orc = Oracle(sim, [])
    rng = np.random.default_rng(2)
    st = reset_state(rng.uniform(0.4, 0.6, (300, 3)))
    st[:, 3:6] = [0.3, -0.2, 0.1]
    out, _ = orc.substep(st, np.zeros((1, 8)), np.zeros((1, 7)))
    # nodes at the edge of the cloud that receive m <= 1e-12 keep v = 0
    # (mpm_simulator.py:233); their vanishing weights leave ~1e-9 in v and,
    # scaled by 4 inv_dx, ~1e-7 in C
    np.testing.assert_allclose(out[:, 3:6], st[:, 3:6], atol=1e-9)
    np.testing.assert_allclose(out[:, 6:15], 0, atol=1e-6)
    np.testing.assert_allclose(out[:, 15:24], st[:, 15:24], atol=1e-12)
    np.testing.assert_allclose(out[:, 0:3], st[:, 0:3] + sim["dt"] * st[:, 3:6], atol=1e-14)


def _grid_node_oracle(gf=1.5, prims=()):
    sim, _ = sim_from_cfg(base_sim(ground_friction=gf))
    return sim, Oracle(sim, list(prims))


def test_grid_op_free_node_and_walls():
    """grid_op (mpm_simulator.py:230-262): v = p/m + dt g 30 away from walls;
    wall band I[d] < 3 zeroes an inward component; upper band I[d] > n - 3
    zeroes an outward one; sticky ground (gf >= 10) zeroes everything."""
    sim, orc = _grid_node_oracle()
    pm = np.array([0.2, -0.1, 0.3, 2.0])
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [30, 30, 30], pm)
    np.testing.assert_allclose(v, pm[:3] / pm[3] + sim["dt"] * np.array(sim["gravity"]) * 30, atol=1e-15)
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [1, 30, 30], np.array([-0.2, 0.1, 0.3, 2.0]))
    assert v[0] == 0 and v[2] == pytest.approx(0.15)
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [62, 30, 30], np.array([0.2, 0.1, 0.3, 2.0]))
    assert v[0] == 0
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [62, 30, 30], np.array([-0.2, 0.1, 0.3, 2.0]))
    assert v[0] < 0
    sim, orc = _grid_node_oracle(gf=20)
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [30, 1, 30], np.array([0.2, -0.1, 0.3, 1.0]))
    np.testing.assert_allclose(v, 0)
    sim, orc = _grid_node_oracle(gf=0)
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [30, 1, 30], np.array([0.2, -0.1, 0.3, 1.0]))
    assert v[1] == 0 and v[0] == pytest.approx(0.2)


def test_grid_op_ground_coulomb_closed_form():
    """Ground friction 0 < gf < 10: tangential velocity scaled by
    max(1 + gf v_n / |v_t|, 0), normal component removed (mpm_simulator.py:247-256)."""
    sim, orc = _grid_node_oracle(gf=1.5)
    p = np.array([0.3, -0.4, 0.4, 1.0])
    v = orc.grid_node(np.zeros((1, 8)), np.zeros((1, 8)), [30, 1, 30], p)
    vin = p[:3] + sim["dt"] * np.array(sim["gravity"]) * 30
    vt = np.array([vin[0], 0, vin[2]])
    scale = max(1 + 1.5 * vin[1] / np.sqrt(vt @ vt + 1e-8), 0)
    np.testing.assert_allclose(v, scale * vt, rtol=1e-12, atol=1e-15)


def test_sdf_closed_forms_and_normals():
    """Sphere / box / capsule SDFs against analytic distances; FD box normal
    (float32 step) against the face normal; rotations applied through inv_trans."""
    rng = np.random.default_rng(4)
    sph = prim_from_cfg(dict(shape="Sphere", radius=0.1))
    box = prim_from_cfg(dict(shape="Box", size=(0.1, 0.05, 0.2)))
    cap = prim_from_cfg(dict(shape="Capsule", h=0.2, r=0.05))
    sim, _ = sim_from_cfg({})
    o = Oracle(sim, [sph, box, cap])
    c = np.array([0.5, 0.4, 0.5])
    ps = np.zeros(8)
    ps[:3] = c
    ps[3] = 1
    x = c + rng.normal(size=(100, 3)) * 0.2
    np.testing.assert_allclose(o.prim_sdf(0, ps, x), np.linalg.norm(x - c, axis=1) - 0.1, atol=1e-12)
    # box: points straight out of the +x face and on a corner diagonal
    xb = c + np.array([[0.3, 0.01, -0.02], [0.2, 0.1, 0.3]])
    d = o.prim_sdf(1, ps, xb)
    assert d[0] == pytest.approx(0.2, abs=1e-6)
    assert d[1] == pytest.approx(np.linalg.norm([0.1, 0.05, 0.1]), abs=1e-6)
    n = o.prim_normal(1, ps, xb[:1])
    np.testing.assert_allclose(n[0], [1, 0, 0], atol=1e-6)
    # capsule: local y is the segment axis, h the full length
    xc = c + np.array([[0.0, 0.35, 0.0], [0.2, 0.05, 0.0]])
    d = o.prim_sdf(2, ps, xc)
    np.testing.assert_allclose(d, [0.35 - 0.1 - 0.05, 0.2 - 0.05], atol=1e-7)
    # rotated box: 90 degrees about z maps local x to world y
    q = np.array([np.cos(np.pi / 4), 0, 0, np.sin(np.pi / 4)])
    ps2 = ps.copy()
    ps2[3:7] = q
    d = o.prim_sdf(1, ps2, c[None] + [[0, 0.3, 0]])
    assert d[0] == pytest.approx(0.2, abs=1e-6)


def _surface_distance(x, surf):
    from scipy.spatial import cKDTree
    return cKDTree(surf).query(x)[0]


def test_cylinder_and_torus_sdfs_and_normals():
    """Cylinder (h the disk radius, r the half height; primitives.py:300-320)
    and Torus (tx the ring radius, ty the tube radius; :336-350): hand-computed
    points, distances of outside points to a densely sampled surface, and the
    analytic normals against the normalised central difference of the SDF,
    in a rotated frame."""
    rng = np.random.default_rng(7)
    cyl = prim_from_cfg(dict(shape="Cylinder", h=0.1, r=0.05))
    tor = prim_from_cfg(dict(shape="Torus", tx=0.2, ty=0.05))
    sim, _ = sim_from_cfg({})
    o = Oracle(sim, [cyl, tor])
    c = np.array([0.5, 0.4, 0.5])
    ps = np.zeros(8)
    ps[:3] = c
    ps[3] = 1
    pts = c + np.array([[0, 0.2, 0], [0.3, 0, 0], [0.2, 0.1, 0], [0.05, 0.01, 0], [0, -0.3, 0]])
    np.testing.assert_allclose(o.prim_sdf(0, ps, pts), [0.15, 0.2, np.hypot(0.1, 0.05), -0.04, 0.25], atol=1e-6)
    np.testing.assert_allclose(o.prim_normal(0, ps, pts[[0, 1, 3, 4]]), [[0, 1, 0], [1, 0, 0], [0, 1, 0], [0, -1, 0]], atol=1e-6)
    pts = c + np.array([[0.2, 0.1, 0], [0.4, 0, 0], [0, 0, -0.2], [0.25, 0, 0]])
    np.testing.assert_allclose(o.prim_sdf(1, ps, pts), [0.05, 0.15, -0.05, 0.0], atol=1e-6)
    np.testing.assert_allclose(o.prim_normal(1, ps, pts[:2]), [[0, 1, 0], [1, 0, 0]], atol=1e-6)
    # sampled surfaces (local frame)
    u = rng.uniform(0, 2 * np.pi, 400000)
    v = rng.uniform(0, 1, 400000)
    side = np.stack([0.1 * np.cos(u), -0.05 + 0.1 * v, 0.1 * np.sin(u)], 1)
    rad = 0.1 * np.sqrt(v)
    caps = np.stack([rad * np.cos(u), np.where(np.arange(len(u)) % 2 == 0, 0.05, -0.05), rad * np.sin(u)], 1)
    cyl_surf = np.concatenate([side, caps])
    w = rng.uniform(0, 2 * np.pi, 400000)
    tor_surf = np.stack([(0.2 + 0.05 * np.cos(w)) * np.cos(u), 0.05 * np.sin(w), (0.2 + 0.05 * np.cos(w)) * np.sin(u)], 1)
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    R = pytest.importorskip("scipy.spatial.transform").Rotation.from_quat([q[1], q[2], q[3], q[0]])
    ps[3:7] = q
    for j, surf, scale in ((0, cyl_surf, 0.25), (1, tor_surf, 0.35)):
        loc = rng.uniform(-scale, scale, (300, 3))
        x = c + R.apply(loc)
        d = o.prim_sdf(j, ps, x)
        out = d > 0.01
        np.testing.assert_allclose(d[out], _surface_distance(loc[out], surf), atol=2e-3)
        # normal = direction of the SDF gradient (central differences in fp64)
        n = o.prim_normal(j, ps, x[out])
        e = 1e-6
        g = np.stack([(o.prim_sdf(j, ps, x[out] + e * np.eye(3)[k]) - o.prim_sdf(j, ps, x[out] - e * np.eye(3)[k])) / (2 * e)
                      for k in range(3)], 1)
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        np.testing.assert_allclose(n, g, atol=1e-5)


def test_forward_kinematics_matches_scipy_rotations():
    """forward_kinematics (primive_base.py:152-156): rot' = w2quat(w) (x) rot, i.e.
    a world-frame rotation by the rotation vector w; pos' = clamp(pos + v)."""
    R = pytest.importorskip("scipy.spatial.transform").Rotation
    rng = np.random.default_rng(5)
    sim, _ = sim_from_cfg({})
    p = prim_from_cfg(dict(shape="Box", lower_bound=(0.1, 0.1, 0.1), upper_bound=(0.9, 0.9, 0.9)))
    g = prim_from_cfg(dict(shape="Gripper", minimal_gap=0.05, maximal_gap=0.3))
    o = Oracle(sim, [p, g])
    for _ in range(20):
        q0 = R.random(random_state=rng.integers(1 << 30))
        w = rng.normal(size=3) * 0.3
        s = np.zeros((2, 8))
        s[:, :3] = rng.uniform(0.05, 0.95, 3)
        qx = q0.as_quat()
        s[:, 3:7] = [qx[3], qx[0], qx[1], qx[2]]
        s[1, 7] = 0.2
        pv = np.zeros((2, 7))
        pv[:, :3] = rng.normal(size=3) * 0.1
        pv[:, 3:6] = w
        pv[1, 6] = rng.normal() * 0.2
        out = o.fk(s, pv)
        ref = (R.from_rotvec(w) * q0).as_quat()
        got = out[0, 3:7]
        assert min(np.abs(got - [ref[3], ref[0], ref[1], ref[2]]).max(), np.abs(got + [ref[3], ref[0], ref[1], ref[2]]).max()) < 1e-12
        np.testing.assert_allclose(out[0, :3], np.clip(s[0, :3] + pv[0, :3], 0.1, 0.9), atol=1e-15)
        # Gripper rotates in its own frame: rot' = rot (x) w2quat(w)
        ref = (q0 * R.from_rotvec(w)).as_quat()
        got = out[1, 3:7]
        assert min(np.abs(got - [ref[3], ref[0], ref[1], ref[2]]).max(), np.abs(got + [ref[3], ref[0], ref[1], ref[2]]).max()) < 1e-12
        assert out[1, 7] == pytest.approx(np.clip(0.2 - pv[1, 6], 0.05, 0.3))


# ---------------------------------------------------------------- adjoint vs finite differences
FD_SHAPES = {
    "Capsule": dict(shape="Capsule", h=0.1, r=0.03, action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Sphere": dict(shape="Sphere", radius=0.04, action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Box": dict(shape="Box", size=(0.03, 0.02, 0.04), action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
    "Knife": dict(shape="Knife", h=(0.05, 0.05), size=(0.03, 0.02, 0.04), prot=(0.9, 0.1, 0.3, 0.2),
                  action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "RollingPin": dict(shape="RollingPin", h=0.1, r=0.03, action=dict(dim=3, scale=(0.5, 0.05, 0.002))),
    "Gripper": dict(shape="Gripper", size=(0.01, 0.03, 0.03), minimal_gap=0.02, init_gap=0.12,
                    action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
    "Chopsticks": dict(shape="Chopsticks", h=0.1, r=0.02, minimal_gap=0.02, init_gap=0.1,
                       action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
    "Cylinder": dict(shape="Cylinder", h=0.04, r=0.02, action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
    "Torus": dict(shape="Torus", tx=0.05, ty=0.015, action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
}


@pytest.mark.parametrize("shape", sorted(FD_SHAPES))
def test_adjoint_matches_central_differences(shape):
    """The hand-written adjoint of whole substeps (p2g, grid_op with collision,
    g2p, primitive kinematics, set_velocity) equals central differences of the
    oracle's own forward pass in fp64."""
    sim, _ = sim_from_cfg(base_sim(n_grid=32, dt=2e-4, substeps=3))
    cfg = dict(FD_SHAPES[shape])
    q = np.array([0.9, 0.2, 0.3, 0.1])
    q /= np.linalg.norm(q)
    cfg.update(init_pos=(0.5, 0.185, 0.5), init_rot=tuple(q))
    o = Oracle(sim, [prim_from_cfg(cfg)])
    rng = np.random.default_rng(0)
    st = reset_state(rng.uniform([0.4, 0.05, 0.4], [0.6, 0.15, 0.6], (300, 3)))
    st[:, 3:6] = rng.normal(0, 0.1, (300, 3))
    ps0 = prim_state_from_cfg(cfg)[None]
    A = cfg["action"]["dim"]
    acts = rng.uniform(-0.5, 0.5, (2, A))
    ro = OracleRollout(o, 3)
    target = np.random.default_rng(2).uniform(0, 1e-4, (32, 32, 32))

    def loss(a):
        fr, _ = ro.run(st, ps0, a)
        m = o.grid_mass(fr[-1])
        return np.abs(m - target).sum() + 1e-3 * (fr[-1][:, :6] ** 2).sum(), fr

    L, fr = loss(acts)
    T = len(fr) - 1
    adj = o.grid_mass_grad(fr[T], np.sign(o.grid_mass(fr[T]) - target))
    adj[:, :6] += 2e-3 * fr[T][:, :6]
    ga, _, _ = ro.grad({T: adj})
    eps = 1e-7
    num = np.zeros_like(acts)
    for s in range(acts.shape[0]):
        for j in range(A):
            a1, a2 = acts.copy(), acts.copy()
            a1[s, j] += eps
            a2[s, j] -= eps
            num[s, j] = (loss(a1)[0] - loss(a2)[0]) / (2 * eps)
    err = np.abs(ga - num).max() / np.abs(num).max()
    assert err < 1e-5, (shape, err, ga, num)


def test_adjoint_state_gradient_matches_central_differences():
    """Adjoint w.r.t. the initial particle state (x, v, C, F) for well separated
    singular values of F."""
    sim, _ = sim_from_cfg(base_sim(n_grid=32, dt=2e-4))
    o = Oracle(sim, [])
    rng = np.random.default_rng(9)
    n = 200
    # away from the wall bands, whose branches FD would straddle
    st = reset_state(rng.uniform([0.4, 0.3, 0.4], [0.6, 0.4, 0.6], (n, 3)))
    st[:, 3:6] = rng.normal(0, 0.1, (n, 3))
    for p in range(n):
        st[p, 15:24] = (_rot(rng) @ np.diag([1.3, 1.0, 0.8]) @ _rot(rng).T).reshape(-1)
    ro = OracleRollout(o, 4)
    w = rng.normal(size=(n, 24))

    def loss(s0):
        fr, _ = ro.run(s0, np.zeros((1, 8)), np.zeros((1, 0)))
        return (fr[-1] * w).sum(), fr

    L, fr = loss(st)
    _, adj0, _ = ro.grad({len(fr) - 1: w})
    for p, c in [(3, 0), (7, 4), (11, 10), (5, 17), (20, 22), (9, 15)]:
        s1, s2 = st.copy(), st.copy()
        s1[p, c] += 1e-6
        s2[p, c] -= 1e-6
        d = (loss(s1)[0] - loss(s2)[0]) / 2e-6
        assert abs(adj0[p, c] - d) < 1e-5 * max(1.0, abs(d)), (p, c, adj0[p, c], d)


@pytest.mark.parametrize("shape", sorted(FD_SHAPES))
def test_min_dist_matches_sdf_and_its_adjoint_matches_central_differences(shape):
    """compute_min_dist (function.py:79-88): one column per primitive (both
    jaws for the two-jaw gripper, whose minimum is Gripper.sdf); its adjoint
    (function.py:154-160) w.r.t. particle positions and the primitive state
    equals central differences in fp64."""
    sim, _ = sim_from_cfg(base_sim(n_grid=32))
    cfg = dict(FD_SHAPES[shape])
    q = np.array([0.9, 0.2, 0.3, 0.1])
    q /= np.linalg.norm(q)
    cfg.update(init_pos=(0.5, 0.185, 0.5), init_rot=tuple(q))
    o = Oracle(sim, [prim_from_cfg(cfg)])
    rng = np.random.default_rng(1)
    st = reset_state(rng.uniform([0.42, 0.12, 0.42], [0.58, 0.26, 0.58], (40, 3)))
    ps = prim_state_from_cfg(cfg)[None]
    d = o.min_dist(st, ps)
    assert d.shape == ((2 if shape == "Gripper" else 1), 40)
    np.testing.assert_allclose(d.min(axis=0), o.prim_sdf(0, ps[0], st[:, :3]), rtol=0, atol=1e-15)
    w = rng.normal(size=d.shape)
    adj, gps = o.min_dist_grad(st, ps, w)
    assert np.all(adj[:, 3:] == 0)
    eps = 1e-7

    def L(s, p):
        return (w * o.min_dist(s, p)).sum()

    for p, c in [(0, 0), (5, 1), (11, 2), (23, 0), (37, 2)]:
        s1, s2 = st.copy(), st.copy()
        s1[p, c] += eps
        s2[p, c] -= eps
        num = (L(s1, ps) - L(s2, ps)) / (2 * eps)
        assert abs(adj[p, c] - num) <= 1e-6 * max(1.0, abs(num)), (p, c, adj[p, c], num)
    for k in range(8):
        p1, p2 = ps.copy(), ps.copy()
        p1[0, k] += eps
        p2[0, k] -= eps
        num = (L(st, p1) - L(st, p2)) / (2 * eps)
        assert abs(gps[0, k] - num) <= 1e-6 * max(1.0, abs(num)), (k, gps[0, k], num)
