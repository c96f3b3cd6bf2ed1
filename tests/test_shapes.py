"""Initial particle clouds from SHAPES (plb/engine/shapes/shape_maker.py):
counts, extents, seeded scatter, rotations against scipy, the global random
state left as it was, and task-file values evaluated without eval."""
import numpy as np
import pytest

from diffskill_amd.engine.shapes import Shapes, _value, quat2mat


def test_box_sphere_capsule_cylinder_counts_and_extents():
    np.random.seed(3)
    before = np.random.get_state()[1].copy()
    sh = Shapes([dict(shape="box", width="(0.2, 0.1, 0.2)", init_pos="(0.5, 0.1, 0.5)"),
                 dict(shape="sphere", radius=0.05, init_pos=(0.3, 0.3, 0.3), color=100),
                 dict(shape="capsule", radius=0.03, height=0.1, init_pos=(0.6, 0.6, 0.6), n_particles=3000),
                 dict(shape="cylinder", radius=0.04, height=0.08, init_pos=(0.2, 0.7, 0.2))])
    assert np.array_equal(np.random.get_state()[1], before)       # generator state restored
    box, sph, cap, cyl = sh.objects
    assert len(box) == int(0.2 * 0.1 * 0.2 / 1e-3 * 30000)     # 30000 per 0.1^3
    assert len(sph) == int(0.05 ** 3 * 4 * np.pi / 3 / 1e-3 * 30000)
    assert len(cap) == 3000 and len(cyl) == int(0.04 ** 3 * 4 * np.pi / 3 / 1e-3 * 30000)
    assert np.all(np.abs(box - [0.5, 0.1, 0.5]) <= [0.1, 0.05, 0.1])
    assert np.all(np.linalg.norm(sph - 0.3, axis=1) <= 0.05 + 1e-12)
    # capsule: axis along z, spherical caps shifted by +-height/2
    d = cap - 0.6
    rho = np.hypot(d[:, 0], d[:, 1])
    assert np.all(rho <= 0.03 + 1e-12) and np.all(np.abs(d[:, 2]) <= 0.05 + 0.03 + 1e-12)
    d = cyl - [0.2, 0.7, 0.2]
    assert np.all(np.hypot(d[:, 0], d[:, 1]) <= 0.04 + 1e-12) and np.all(np.abs(d[:, 2]) <= 0.04 + 1e-12)
    x, colors = sh.get()
    assert x.shape == (len(box) + len(sph) + 3000 + len(cyl), 3) and len(colors) == len(x)
    assert colors[0] == (127 << 16) + 127 and colors[len(box)] == 100
    # the same global state gives the same particles
    np.random.seed(3)
    x2, _ = Shapes([dict(shape="box", width="(0.2, 0.1, 0.2)", init_pos="(0.5, 0.1, 0.5)"),
                    dict(shape="sphere", radius=0.05, init_pos=(0.3, 0.3, 0.3), color=100),
                    dict(shape="capsule", radius=0.03, height=0.1, init_pos=(0.6, 0.6, 0.6), n_particles=3000),
                    dict(shape="cylinder", radius=0.04, height=0.08, init_pos=(0.2, 0.7, 0.2))]).get()
    np.testing.assert_array_equal(x, x2)


def test_scatter_is_seeded_and_multi_shapes_split_the_count():
    """gather_move.yml's SHAPES: 40 seeded centres x 50 noisy particles."""
    cfg = [dict(shape="scatter", pos_min="(0.64, 0.02, 0.38)", pos_max="(0.76, 0.035, 0.62)",
                color="(0.7568, 0.6039, 0.4196)", seed=0)]
    np.random.seed(1)
    a = Shapes(cfg).objects[0]
    np.random.seed(2)
    b = Shapes(cfg).objects[0]
    np.testing.assert_array_equal(a, b)
    assert a.shape == (2000, 3)
    # heights: the centre's y plus 0.2 x N(0, 0.004) + 0.1
    assert abs(a[:, 1].mean() - (0.0275 + 0.1)) < 2e-3
    ms = Shapes([dict(shape="multibox", all_pos=[(0.3, 0.1, 0.3), (0.6, 0.1, 0.6)], all_width=[(0.1, 0.1, 0.1), (0.2, 0.1, 0.1)],
                      all_rot=[None, None], color=5)])
    n = min(30000, int(0.003 / 1e-3 * 30000))
    assert [len(o) for o in ms.objects] == [int(n * 0.001 / 0.003), int(n * 0.002 / 0.003)]


def test_rotation_about_the_mean_matches_scipy():
    R = pytest.importorskip("scipy.spatial.transform").Rotation
    rng = np.random.default_rng(0)
    for _ in range(10):
        q = rng.normal(size=4) * rng.uniform(0.3, 3)           # not unit: normalised implicitly
        np.testing.assert_allclose(quat2mat(q), R.from_quat([q[1], q[2], q[3], q[0]]).as_matrix(), atol=1e-12)
    np.random.seed(0)
    plain = Shapes([dict(shape="box", width=(0.2, 0.02, 0.02), init_pos=(0.5, 0.5, 0.5))]).objects[0]
    np.random.seed(0)
    q = (np.cos(np.pi / 4), 0.0, np.sin(np.pi / 4), 0.0)       # 90 degrees about y
    rot = Shapes([dict(shape="box", width=(0.2, 0.02, 0.02), init_pos=(0.5, 0.5, 0.5), init_rot=q)]).objects[0]
    m = plain.mean(axis=0)
    np.testing.assert_allclose(rot, (plain - m) @ R.from_quat([0, q[2], 0, q[0]]).as_matrix().T + m, atol=1e-12)
    assert np.ptp(rot[:, 2]) > 0.19 and np.ptp(rot[:, 0]) < 0.021


def test_task_file_values_are_evaluated_without_eval():
    assert _value("((150<<8)+ (150<<16))") == (150 << 8) + (150 << 16)
    assert _value("(0.05)") == 0.05 and _value("(0.5, -0.1, 2*0.25)") == (0.5, -0.1, 0.5)
    with pytest.raises(ValueError):
        _value("__import__('os').system('true')")
    with pytest.raises(NotImplementedError):
        Shapes([dict(shape="torus_cloud")])
