"""Independent pins of the oracle functions that restate reference text with
no library routine to compare against: the constitutive stress, the APIC
transfer, the tool collision response, the kinematics of the rolling pins,
chopsticks and two-capsule gripper, and the prism / knife / two-capsule
SDFs.  Each check comes from the mathematics (energy derivatives by central
differences, exact reproduction of linear fields, closed-form contact
responses, scipy rotation composition, distances to known faces), so that a
dropped term, a wrong sign or scale, or a transposed operand in the oracle
fails it."""
import numpy as np
import pytest

from oracle import Oracle, prim_from_cfg, prim_state_from_cfg, reset_state, sim_from_cfg

Rot = pytest.importorskip("scipy.spatial.transform").Rotation


def wxyz(r):
    q = r.as_quat()
    return np.array([q[3], q[0], q[1], q[2]])


def same_rotation(q, r):
    ref = wxyz(r)
    return min(np.abs(q - ref).max(), np.abs(q + ref).max())


# ---------------------------------------------------------------- stress
def energy(F, mu, lam):
    """Fixed-corotated energy psi = mu sum (sigma - 1)^2 + lam / 2 (J - 1)^2."""
    s = np.linalg.svd(F, compute_uv=False)
    J = np.linalg.det(F)
    return mu * ((s - 1) ** 2).sum() + 0.5 * lam * (J - 1) ** 2


def energy_grad(F, mu, lam, h=1e-6):
    P = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            E = np.zeros((3, 3))
            E[i, j] = h
            P[i, j] = (energy(F + E, mu, lam) - energy(F - E, mu, lam)) / (2 * h)
    return P


@pytest.mark.parametrize("case", ["elastic", "yield", "with_C"])
def test_stress_is_energy_derivative_times_Ft(case):
    """p2g's stress (mpm_simulator.py:205-215) is -dt p_vol 4 inv_dx^2
    (d psi / dF)(F') F'^T at the post-return-map F', psi the fixed-corotated
    energy: checked against central differences of psi computed from numpy's
    SVD and determinant, and the returned F' against F_tmp = (I + dt C) F
    when the particle stays elastic."""
    sim, _ = sim_from_cfg(dict(E=5e3, nu=0.2, yield_stress=50.0, gravity=(0, 0, 0)))
    o = Oracle(sim, [])
    rng = np.random.default_rng(11)
    mu, lam = sim["mu"], sim["lam"]
    stress_k = -sim["dt"] * sim["p_vol"] * 4 * sim["inv_dx"] ** 2
    for _ in range(6):
        Q = Rot.random(random_state=rng.integers(1 << 30)).as_matrix()
        W = Rot.random(random_state=rng.integers(1 << 30)).as_matrix()
        if case == "yield":
            s = np.array([1.25, 0.9, 0.8]) * rng.uniform(0.95, 1.05, 3)
        else:
            s = 1 + rng.uniform(-4e-3, 4e-3, 3)
        F = Q @ np.diag(s) @ W.T
        st = reset_state(np.array([[0.5, 0.5, 0.5]]))
        st[0, 15:24] = F.reshape(-1)
        C = np.zeros((3, 3))
        if case == "with_C":
            C = rng.normal(size=(3, 3)) * 3
            st[0, 6:15] = C.reshape(-1)
        Fn, stress = o.particle_stress(st)
        Fn = Fn.reshape(3, 3)
        Ftmp = (np.eye(3) + sim["dt"] * C) @ F
        if case == "yield":
            assert np.abs(Fn - Ftmp).max() > 1e-3          # the return map moved F
        else:
            np.testing.assert_allclose(Fn, Ftmp, atol=1e-14)
        expect = stress_k * energy_grad(Fn, mu, lam) @ Fn.T
        np.testing.assert_allclose(stress.reshape(3, 3), expect, rtol=1e-6, atol=1e-8 * np.abs(expect).max())


# ---------------------------------------------------------------- APIC
def test_apic_transfer_reproduces_a_linear_velocity_field():
    """With no stress (E = 0), no gravity, no tools and no walls in reach, a
    velocity field v = A x + b carried with C = A is reproduced exactly by
    p2g -> grid_op -> g2p (mpm_simulator.py:198-283): the grid holds
    v_i = A x_i + b (the affine term p_mass C (x_i - x_p) completes the
    field), and g2p returns v = A x_p + b (B-spline weights reproduce linear
    functions) and C = 4 inv_dx^2 sum w v_i (x_i - x_p)^T = A (their second
    moment is dx^2 / 4 I).  A wrong APIC scale, a dropped affine term or a
    transposed C fails it."""
    sim, _ = sim_from_cfg(dict(E=0.0, nu=0.2, yield_stress=50.0, gravity=(0, 0, 0), n_grid=32))
    o = Oracle(sim, [])
    rng = np.random.default_rng(12)
    dx = sim["dx"]
    # particles whose fractional cell offsets stay in [0.6, 1.4]: every node
    # of every stencil gets a weight of at least 0.5 * 0.1^2
    cells = rng.integers(12, 20, (400, 3))
    x = (cells + rng.uniform(0.6, 1.4, (400, 3))) * dx
    A = rng.normal(size=(3, 3))
    b = rng.normal(size=3)
    st = reset_state(x)
    st[:, 3:6] = x @ A.T + b
    st[:, 6:15] = A.reshape(-1)
    out, _ = o.substep(st, np.zeros((1, 8)), np.zeros((1, 7)))
    np.testing.assert_allclose(out[:, 3:6], x @ A.T + b, atol=1e-11)
    np.testing.assert_allclose(out[:, 6:15], np.tile(A.reshape(-1), (400, 1)), atol=1e-9)
    np.testing.assert_allclose(out[:, 0:3], x + sim["dt"] * (x @ A.T + b), atol=1e-14)
    np.testing.assert_allclose(out[:, 15:24], np.tile((np.eye(3) + sim["dt"] * A).reshape(-1), (400, 1)), atol=1e-14)


# ---------------------------------------------------------------- collide
def _sphere_node(friction, r=0.1, softness=666.0):
    sim, _ = sim_from_cfg(dict(gravity=(0, 0, 0)))
    p = prim_from_cfg(dict(shape="Sphere", radius=r, friction=friction), softness)
    return sim, Oracle(sim, [p])


def _collide(o, c0, c1, I, v, q0=(1, 0, 0, 0), q1=None):
    ps0 = np.zeros((1, 8))
    ps0[0, :3] = c0
    ps0[0, 3:7] = q0
    ps1 = ps0.copy()
    ps1[0, :3] = c1
    if q1 is not None:
        ps1[0, 3:7] = q1
    return o.grid_node(ps0, ps1, I, np.array([v[0], v[1], v[2], 1.0]))


def test_collide_closed_forms():
    """Primitive.collide (primive_base.py:96-120) against its contact law at a
    node of a static sphere (analytic normal n = (x - c) / |x - c|):
    - far from the tool (influence exp(-666 d) < 0.1, d > 0): v unchanged;
    - inside, approaching, frictionless: the normal part removed;
    - inside, approaching, friction mu: the tangential part scaled by
      max(0, |v_t| + mu v.n) / |v_t|, both with the base length's 1e-8;
    - inside, approaching, high friction: stuck (v = 0);
    - inside, separating: v unchanged;
    - just outside (0.1 < influence < 1): the influence-weighted blend."""
    I = np.array([32, 32, 32])
    x = I / 64.0
    nrm = np.array([1.0, 2.0, -0.5])
    nrm /= np.linalg.norm(nrm)
    v_in = np.array([-0.3, 0.1, 0.2])   # v . n < 0: approaching the tool's surface from outside
    assert v_in @ nrm < 0
    L = lambda a: np.sqrt(a @ a + 1e-8)   # noqa: E731  utils.length

    def centre(d, r=0.1):  # sphere centre putting x at signed distance d along +n
        return x - (r + d) * nrm

    for fr in (0.0, 0.9, 50.0):
        sim, o = _sphere_node(fr)
        # far field
        c = centre(0.01)
        np.testing.assert_allclose(_collide(o, c, c, I, v_in), v_in, atol=1e-15)
        # inside: influence 1
        c = centre(-0.002)
        vn = v_in @ nrm
        vt = v_in - vn * nrm
        expect = vt / L(vt) * max(0.0, L(vt) + vn * fr)
        np.testing.assert_allclose(_collide(o, c, c, I, v_in), expect, atol=1e-14)
        # separating
        np.testing.assert_allclose(_collide(o, c, c, I, -v_in), -v_in, atol=1e-15)
        # just outside: influence between 0.1 and 1
        d = 0.002
        c = centre(d)
        infl = np.exp(-d * 666.0)
        assert 0.1 < infl < 1
        expect_t = vt / L(vt) * max(0.0, L(vt) + vn * fr)
        np.testing.assert_allclose(_collide(o, c, c, I, v_in), v_in * (1 - infl) + expect_t * infl, atol=1e-14)
    # stuck at high friction
    sim, o = _sphere_node(50.0)
    c = centre(-0.002)
    np.testing.assert_allclose(_collide(o, c, c, I, v_in), 0, atol=1e-15)


def test_collider_velocity_of_moving_tools():
    """collider_v (primive_base.py:87-94): the velocity the tool imparts at a
    node, (p_{f+1} + R_{f+1} R_f^-1 (x - p_f) - x) / dt; a stuck contact
    (high friction) takes it exactly -- for a pure translation (dp / dt) and
    for a rotation about the tool's centre (scipy rotation of the offset)."""
    I = np.array([32, 32, 32])
    x = I / 64.0
    sim, o = _sphere_node(50.0)
    dt = sim["dt"]
    c0 = x - np.array([0.09, 0.0, 0.0])          # x 0.01 inside the sphere
    dp = np.array([1e-4, -2e-4, 5e-5])
    v_in = np.array([-0.5, 0.02, 0.01])          # pushing into the moving tool
    np.testing.assert_allclose(_collide(o, c0, c0 + dp, I, v_in), dp / dt, rtol=1e-9)
    rng = np.random.default_rng(13)
    r0 = Rot.random(random_state=7)
    dr = Rot.from_rotvec(rng.normal(size=3) * 1e-3)
    r1 = dr * r0
    vc = (dr.apply(x - c0) + c0 - x) / dt
    v = _collide(o, c0, c0, I, vc - 0.5 * (x - c0) / np.linalg.norm(x - c0), q0=wxyz(r0), q1=wxyz(r1))
    np.testing.assert_allclose(v, vc, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- kinematics
def _fk_case(shape, **cfg):
    sim, _ = sim_from_cfg({})
    p = prim_from_cfg(dict(shape=shape, lower_bound=(0.1, 0.1, 0.1), upper_bound=(0.9, 0.9, 0.9), **cfg))
    return Oracle(sim, [p])


@pytest.mark.parametrize("shape", ["RollingPin", "RollingPinExt"])
def test_rolling_pin_kinematics_matches_scipy(shape):
    """RollingPin / RollingPinExt.forward_kinematics (primitives.py:103-136):
    the pin rolls about its own axis by dw and turns about the world y by
    -dtheta, R' = R_y(-dtheta) R R_y(dw); it moves along (0, 1, 0) x
    (R (0, -1, 0)) by 0.03 dw (+ w[0] for the Ext variant) and to height
    change dy."""
    rng = np.random.default_rng(14)
    o = _fk_case(shape, h=0.2, r=0.03)
    for _ in range(20):
        r0 = Rot.random(random_state=rng.integers(1 << 30))
        s = np.zeros(8)
        s[:3] = rng.uniform(0.3, 0.7, 3)
        s[3:7] = wxyz(r0)
        pv = np.zeros(7)
        pv[:3] = rng.normal(size=3) * [0.3, 0.2, 0.01]
        pv[3:6] = rng.normal(size=3) * 0.01
        out = o.fk(s[None], pv[None])[0]
        dw, dth, dy = pv[:3]
        ref = Rot.from_rotvec([0, -dth, 0]) * r0 * Rot.from_rotvec([0, dw, 0])
        assert same_rotation(out[3:7], ref) < 1e-12
        y_dir = r0.apply([0, -1, 0])
        k = 0.03 * dw + (pv[3] if shape == "RollingPinExt" else 0.0)
        move = np.cross([0, 1, 0], y_dir) * k
        move[1] = dy
        np.testing.assert_allclose(out[:3], np.clip(s[:3] + move, 0.1, 0.9), atol=1e-15)


@pytest.mark.parametrize("shape", ["Chopsticks", "Gripper2"])
def test_two_finger_kinematics_matches_scipy(shape):
    """Chopsticks / Gripper2.forward_kinematics (primitives.py:230-235,
    588-592): R' = R R(w) (a rotation in the tool's own frame), p' =
    clamp(p + v), gap' = max(gap - gap_vel, minimal_gap), and for Gripper2
    also at most maximal_gap."""
    rng = np.random.default_rng(15)
    o = _fk_case(shape, h=0.1, r=0.02, minimal_gap=0.03, maximal_gap=0.25)
    for _ in range(20):
        r0 = Rot.random(random_state=rng.integers(1 << 30))
        s = np.zeros(8)
        s[:3] = rng.uniform(0.3, 0.7, 3)
        s[3:7] = wxyz(r0)
        s[7] = rng.uniform(0.05, 0.2)
        pv = np.zeros(7)
        pv[:3] = rng.normal(size=3) * 0.05
        pv[3:6] = rng.normal(size=3) * 0.3
        pv[6] = rng.normal() * 0.2
        out = o.fk(s[None], pv[None])[0]
        assert same_rotation(out[3:7], r0 * Rot.from_rotvec(pv[3:6])) < 1e-12
        np.testing.assert_allclose(out[:3], np.clip(s[:3] + pv[:3], 0.1, 0.9), atol=1e-15)
        g = max(s[7] - pv[6], 0.03)
        if shape == "Gripper2":
            g = min(g, 0.25)
        assert out[7] == pytest.approx(g, abs=1e-15)


# ---------------------------------------------------------------- SDFs
def _posed(shape, rot, **cfg):
    sim, _ = sim_from_cfg({})
    o = Oracle(sim, [prim_from_cfg(dict(shape=shape, **cfg))])
    ps = np.zeros(8)
    ps[:3] = [0.5, 0.4, 0.5]
    ps[3:7] = wxyz(rot)
    return o, ps


def test_prism_and_knife_sdfs_at_face_points():
    """Prism._sdf (primitives.py:711-718): a triangular prism, apex up, edge
    along local z (half length h[1]), faces at y = -h[0] / 2 (bottom) and
    0.866 |x| + 0.5 y = h[0] / 2 (the slanted sides); points straight out of a
    face are at the SDF value of their distance to it, and the FD normal
    there is the face normal; the frame is the prism rotation prot composed
    with the tool rotation.  Knife = max(prism, box) (primitives.py:769-773):
    outside the box's thin z extent the box face decides."""
    rng = np.random.default_rng(16)
    h0, h1 = 0.08, 0.1
    prot = Rot.from_rotvec([0.3, -0.2, 0.5])
    r = Rot.random(random_state=3)
    frame = r * prot          # local prism coordinates -> world offsets
    o, ps = _posed("Prism", r, h=(h0, h1), prot=tuple(wxyz(prot)))
    n_side = np.array([0.866025, 0.5, 0.0])
    n_side_unit = n_side / np.linalg.norm(n_side)
    for t in (0.01, 0.03):
        # bottom face (local -y), slanted face (+x side), end face (+z)
        local = np.array([[0.004, -h0 / 2 - t, 0.01],
                          [0.0, 0.0, 0.02] + n_side_unit * (h0 / 2 / np.linalg.norm(n_side) + t),
                          [0.0, -0.01, h1 + t]])
        x = ps[:3] + frame.apply(local)
        d = o.prim_sdf(0, ps, x)
        # the bound SDF measures the slanted face along its unnormalised
        # normal (|(0.866025, 0.5)| = 0.99999993)
        np.testing.assert_allclose(d, [t, t * np.linalg.norm(n_side), t], atol=1e-9)
        nrm = o.prim_normal(0, ps, x)
        np.testing.assert_allclose(nrm, frame.apply(np.array([[0, -1, 0], n_side_unit, [0, 0, 1]])), atol=2e-6)
    # knife: thin box (half z 0.002) cut by the same prism
    o, ps = _posed("Knife", r, h=(h0, h1), size=(0.1, 0.05, 0.002), prot=tuple(wxyz(prot)))
    t = 0.02
    local_box = np.array([0.0, -0.01, 0.002 + t])      # straight out of the box's z face (box frame = tool frame)
    x = ps[:3] + r.apply(local_box)
    assert o.prim_sdf(0, ps, x[None])[0] == pytest.approx(t, abs=1e-9)
    # a point the box contains deep inside, outside the prism: the prism decides
    loc = np.array([0.0, -h0 / 2 - 0.01, 0.0])
    x = ps[:3] + frame.apply(loc)
    assert np.abs(r.inv().apply(x - ps[:3])).max() < 0.05
    assert o.prim_sdf(0, ps, x[None])[0] == pytest.approx(0.01, abs=1e-9)


def test_two_capsule_sdfs_at_known_points():
    """Chopsticks._sdf (primitives.py:249-254): capsules of radius r along
    local y in [-h, 0] at local x = -gap/2 and +gap/2; Gripper2.sdf
    (:611-620): capsules of radius r along local y in [-h/2, h/2] centred at
    position + R (+-gap/2, 0, 0).  Points at known distances: beside a
    finger, between the fingers, beyond a segment end."""
    r = Rot.from_rotvec([0.4, 0.1, -0.3])
    h, rad, gap = 0.1, 0.02, 0.08
    o, ps = _posed("Chopsticks", r, h=h, r=rad, minimal_gap=0.01)
    ps[7] = gap
    local = np.array([[gap / 2 + rad + 0.03, -h / 2, 0.0],     # beside the +x finger
                      [0.0, -h / 2, 0.0],                       # between them
                      [-gap / 2, 0.05, 0.0],                    # above the -x finger's top end (y = 0)
                      [gap / 2, -h - 0.04, 0.0]])               # below the +x finger's bottom end
    x = ps[:3] + r.apply(local)
    np.testing.assert_allclose(o.prim_sdf(0, ps, x), [0.03, gap / 2 - rad, 0.05 - rad, 0.04 - rad], atol=1e-9)
    o, ps = _posed("Gripper2", r, h=h, r=rad, minimal_gap=0.01, maximal_gap=1.0)
    ps[7] = gap
    local = np.array([[gap / 2 + rad + 0.03, 0.0, 0.0],
                      [0.0, 0.02, 0.0],
                      [-gap / 2, h / 2 + 0.05, 0.0],
                      [gap / 2, -h / 2 - 0.04, 0.01]])
    x = ps[:3] + r.apply(local)
    np.testing.assert_allclose(o.prim_sdf(0, ps, x), [0.03, gap / 2 - rad, 0.05 - rad, np.hypot(0.04, 0.01) - rad],
                               atol=1e-9)
