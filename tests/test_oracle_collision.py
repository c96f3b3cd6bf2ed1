"""Pins of the oracle's tool-tool collision projection (mpm_simulator.py:286-305,
primive_base.py:135-150,284-287): surface points and collision indices
against scipy rotations and numpy's argmin, the projected pose against the
closed form of an axis-aligned box contact, and the adjoint against central
differences of the oracle's forward (the pass alone and whole rollouts)."""
import numpy as np
import pytest

from oracle import Oracle, OracleRollout, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg

N_POINTS = 600  # num_particle_collision (mpm_simulator.py:59)


def rand_num(n_pairs, n_points=N_POINTS):
    """MPMSimulator.initialize (mpm_simulator.py:87-96): np.random.seed(42);
    uniform(-1.5, 1.5, (pairs, 600, 3)), stored in a float32 field."""
    return np.random.RandomState(42).uniform(-1.5, 1.5, (n_pairs, n_points, 3)).astype(np.float32).astype(np.float64)


def unit_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


OBSTACLE = dict(shape="Box", size=(0.2, 0.05, 0.15))
TOOLS = {
    "Box": dict(shape="Box", size=(0.07, 0.02, 0.05), action=dict(dim=6, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05))),
    "Gripper": dict(shape="Gripper", size=(0.015, 0.04, 0.03), minimal_gap=0.02, init_gap=0.1,
                    action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
    "Sphere": dict(shape="Sphere", radius=0.04, action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Capsule": dict(shape="Capsule", h=0.1, r=0.03, action=dict(dim=3, scale=(0.01, 0.01, 0.01))),
    "Gripper2": dict(shape="Gripper2", h=0.06, r=0.015, minimal_gap=0.02, init_gap=0.1,
                     action=dict(dim=7, scale=(0.01, 0.01, 0.01, 0.05, 0.05, 0.05, 0.01))),
}


def overlapping_states(rng, tool_shapes):
    """Primitive states: the tools (rotated) straddling the obstacle's top face,
    the obstacle last (rotated too)."""
    ps = []
    for sh in tool_shapes:
        s = prim_state_from_cfg(dict(TOOLS[sh]))
        s[:3] = [0.5 + rng.uniform(-0.05, 0.05), 0.36 + rng.uniform(-0.01, 0.01), 0.5 + rng.uniform(-0.05, 0.05)]
        s[3:7] = unit_quat(rng)
        ps.append(s)
    s = prim_state_from_cfg(dict(OBSTACLE))
    s[:3] = [0.5, 0.3, 0.5]
    q = np.array([1.0, 0.05, 0.0, 0.03])
    s[3:7] = q / np.linalg.norm(q)
    ps.append(s)
    return np.stack(ps)


def make_oracle(tool_shapes, pairs):
    sim, _ = sim_from_cfg(dict(n_grid=32, dt=2e-4, substeps=3))
    prims = [prim_from_cfg(dict(TOOLS[sh])) for sh in tool_shapes] + [prim_from_cfg(dict(OBSTACLE))]
    return Oracle(sim, prims, collisions=(pairs, rand_num(len(pairs))))


def test_surface_points_and_collision_index_against_scipy_and_argmin():
    """Point k of pair (i, j) is qrot(rot_j, clip(rand_k * size_j, -size_j,
    size_j)) + pos_j; collision_idx is the first index of the smallest
    negative SDF of i over those points (numpy argmin), -1 if none."""
    R = pytest.importorskip("scipy.spatial.transform").Rotation
    rng = np.random.default_rng(4)
    shapes = ["Box", "Gripper", "Sphere"]
    pairs = [(0, 3), (1, 3), (2, 3)]
    o = make_oracle(shapes, pairs)
    rand = rand_num(3)
    for trial in range(4):
        ps = overlapping_states(rng, shapes)
        if trial == 3:
            ps[2, :3] = [0.9, 0.8, 0.9]           # sphere far away: no contact
        out, rec = o.tool_collide(ps)
        size_j = np.array(OBSTACLE["size"])
        qj = ps[3, 3:7]
        rot = R.from_quat([qj[1], qj[2], qj[3], qj[0]])
        for p, (i, j) in enumerate(pairs):
            w = rot.apply(np.clip(rand[p] * size_j, -size_j, size_j)) + ps[j, :3]
            d = o.prim_sdf(i, ps[i], w)
            k = int(np.argmin(d)) if d.min() < 0 else -1
            assert rec["idx"][p] == k, (trial, p, rec["idx"][p], k)
            if k >= 0:
                np.testing.assert_allclose(rec["pt"][p], w[k], rtol=0, atol=1e-14)
                n = o.prim_normal(i, ps[i], w[k][None])[0]
                np.testing.assert_allclose(out[i, :3], ps[i, :3] + n / np.linalg.norm(n) * d[k], rtol=0, atol=1e-14)
            else:
                np.testing.assert_array_equal(out[i], ps[i])
            np.testing.assert_array_equal(out[i, 3:], ps[i, 3:])
        np.testing.assert_array_equal(out[3], ps[3])
        if trial == 3:
            assert rec["idx"][2] == -1


def test_projection_puts_the_deepest_point_on_an_axis_aligned_box_face():
    """Closed form: an unrotated box overlapping an unrotated box obstacle is
    moved along one axis by exactly the penetration of the deepest sampled
    point, which then lies on the box's surface (SDF 0)."""
    rng = np.random.default_rng(8)
    o = make_oracle(["Box"], [(0, 1)])
    s_tool = np.array(TOOLS["Box"]["size"])
    for _ in range(10):
        ps = np.zeros((2, 8))
        ps[:, 3] = 1.0
        ps[1, :3] = [0.5, 0.3, 0.5]
        ps[0, :3] = ps[1, :3] + [rng.uniform(-0.1, 0.1), rng.uniform(0.04, 0.065), rng.uniform(-0.1, 0.1)]
        out, rec = o.tool_collide(ps)
        assert rec["idx"][0] >= 0
        x = rec["pt"][0]
        shift = out[0, :3] - ps[0, :3]
        ax = int(np.argmax(np.abs(shift)))
        assert np.abs(np.delete(shift, ax)).max() < 1e-12
        # the point's depth below the face the shift is normal to, less the
        # sqrt(1e-14) that length() adds to every box SDF (primitives.py:14-15)
        depth = s_tool[ax] - abs(x[ax] - ps[0, ax])
        assert abs(abs(shift[ax]) - (depth - 1e-7)) < 1e-12
        assert abs(o.prim_sdf(0, out[0], x[None])[0]) < 1e-9


@pytest.mark.parametrize("shape", sorted(TOOLS))
def test_collision_adjoint_matches_central_differences(shape):
    """d/d(state of i) of w . (projected states) equals the adjoint of the
    pass (position, rotation, gap of tool i; the obstacle's points fixed)."""
    rng = np.random.default_rng(sorted(TOOLS).index(shape))
    o = make_oracle([shape], [(0, 1)])
    for trial in range(3):
        ps = overlapping_states(rng, [shape])
        out, rec = o.tool_collide(ps)
        if rec["idx"][0] < 0:
            continue
        w = rng.normal(size=ps.shape)
        g = o.tool_collide_rev(out, rec, w)
        eps = 1e-7
        comps = range(8) if shape in ("Gripper", "Gripper2") else range(7)
        for c in comps:
            p1, p2 = ps.copy(), ps.copy()
            p1[0, c] += eps
            p2[0, c] -= eps
            o1, r1 = o.tool_collide(p1)
            o2, r2 = o.tool_collide(p2)
            assert r1["idx"][0] == r2["idx"][0] == rec["idx"][0]
            num = ((o1 - o2) * w).sum() / (2 * eps)
            assert abs(g[0, c] - num) < 1e-6 * max(1.0, abs(num)), (shape, trial, c, g[0, c], num)
        # the obstacle's state gets nothing: its points carry no gradient
        np.testing.assert_array_equal(g[1], w[1])


def test_two_pairs_of_one_tool_compose_in_order():
    """A tool in two pairs is projected by the first pair, then the second
    pair projects the moved tool; the adjoint composes in reverse."""
    rng = np.random.default_rng(21)
    sim, _ = sim_from_cfg(dict(n_grid=32, dt=2e-4, substeps=3))
    prims = [prim_from_cfg(dict(TOOLS["Box"])), prim_from_cfg(dict(OBSTACLE)), prim_from_cfg(dict(OBSTACLE))]
    o = Oracle(sim, prims, collisions=([(0, 1), (0, 2)], rand_num(2)))
    ps = np.zeros((3, 8))
    ps[:, 3] = 1.0
    ps[0, :3] = [0.5, 0.36, 0.5]
    ps[0, 3:7] = unit_quat(rng)
    ps[1, :3] = [0.5, 0.3, 0.5]
    ps[2, :3] = [0.5, 0.42, 0.55]
    out, rec = o.tool_collide(ps)
    assert (rec["idx"] >= 0).all()
    np.testing.assert_allclose(rec["pre"][0], ps[0, :3], rtol=0, atol=0)
    assert np.abs(rec["pre"][1] - ps[0, :3]).max() > 1e-4     # second pair starts from the first's result
    w = rng.normal(size=ps.shape)
    g = o.tool_collide_rev(out, rec, w)
    for c in range(7):
        p1, p2 = ps.copy(), ps.copy()
        p1[0, c] += 1e-7
        p2[0, c] -= 1e-7
        num = ((o.tool_collide(p1)[0] - o.tool_collide(p2)[0]) * w).sum() / 2e-7
        assert abs(g[0, c] - num) < 1e-6 * max(1.0, abs(num)), (c, g[0, c], num)


def test_rollout_with_collision_action_gradient_matches_central_differences():
    """Whole substeps with the projection between forward kinematics and the
    grid update: the adjoint w.r.t. the actions of a box tool pushed into a
    box obstacle over a dough block equals central differences in fp64."""
    sim, _ = sim_from_cfg(dict(E=5e3, nu=0.2, yield_stress=50.0, gravity=(0, -9.8, 0), ground_friction=1.5,
                               n_grid=32, dt=2e-4, substeps=3))
    tool = dict(TOOLS["Box"], init_pos=(0.5, 0.2, 0.5))
    q = np.array([0.95, 0.1, 0.2, 0.05])
    tool["init_rot"] = tuple(q / np.linalg.norm(q))
    obst = dict(OBSTACLE, init_pos=(0.5, 0.14, 0.62), init_rot=(1.0, 0.0, 0.0, 0.0))
    o = Oracle(sim, [prim_from_cfg(tool), prim_from_cfg(obst)], collisions=([(0, 1)], rand_num(1)))
    rng = np.random.default_rng(3)
    st = reset_state(rng.uniform([0.4, 0.05, 0.4], [0.6, 0.12, 0.6], (300, 3)))
    st[:, 3:6] = rng.normal(0, 0.1, (300, 3))
    ps0 = np.stack([prim_state_from_cfg(tool), prim_state_from_cfg(obst)])
    acts = rng.uniform(-0.5, 0.5, (2, 6))
    acts[:, 1] = -0.8                       # pushed down into the obstacle
    ro = OracleRollout(o, 3)
    w = rng.normal(size=(300, 6))
    wp = rng.normal(size=(2, 8))

    def loss(a):
        fr, ps = ro.run(st, ps0, a)
        return (w * fr[-1][:, :6]).sum() + (wp * ps[-1]).sum(), fr, ps

    L, fr, ps = loss(acts)
    assert all(c["idx"][0] >= 0 for c in ro.colls), "the tool must touch the obstacle in every substep"
    T = len(fr) - 1
    adj = np.zeros_like(fr[T])
    adj[:, :6] = w
    ga, _, _ = ro.grad({T: adj}, {T: wp})
    eps = 1e-7
    num = np.zeros_like(acts)
    for s in range(2):
        for j in range(6):
            a1, a2 = acts.copy(), acts.copy()
            a1[s, j] += eps
            a2[s, j] -= eps
            num[s, j] = (loss(a1)[0] - loss(a2)[0]) / (2 * eps)
    err = np.abs(ga - num).max() / np.abs(num).max()
    assert err < 1e-5, (err, ga, num)
