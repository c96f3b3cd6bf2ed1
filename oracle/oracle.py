"""ctypes front-end of the CPU oracle (``mpm_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py``, always as the checker.  The
product package never imports it.

Scenes are described with plain dicts so this module does not depend on the
product package:

``sim``   n_grid, dx, inv_dx, dt, p_vol, p_mass, mu, lam, yield_stress,
          gravity (3), ground_friction, lower_bound
``prim``  type (int code, see ``PRIM_TYPES``), action_dim, friction, softness,
          size (3), h, r, ph (2), prot (4), lo (3), hi (3), min_gap, max_gap,
          action_scale (<= 7)

Particle state per frame is an ``(n, 24)`` array (x, v, C row-major,
F row-major); primitive state per frame ``(n_prims, 8)`` (pos, quat wxyz,
gap); per-substep primitive velocity ``(n_prims, 7)`` (v, w, gap_vel).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

PRIM_TYPES = {
    "Sphere": 0, "Capsule": 1, "RollingPin": 2, "RollingPinExt": 3, "Box": 4,
    "Gripper": 5, "Prism": 6, "Knife": 7, "Chopsticks": 8, "Gripper2": 9, "Cylinder": 10, "Torus": 11,
}


class OrcSim(ctypes.Structure):
    _fields_ = [
        ("n_grid", ctypes.c_int), ("n_prims", ctypes.c_int),
        ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("p_vol", ctypes.c_double), ("p_mass", ctypes.c_double),
        ("mu", ctypes.c_double), ("lam", ctypes.c_double), ("yield_stress", ctypes.c_double),
        ("gravity", ctypes.c_double * 3), ("ground_friction", ctypes.c_double),
        ("lower_bound", ctypes.c_double),
    ]


class OrcPrim(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int), ("action_dim", ctypes.c_int),
        ("friction", ctypes.c_double), ("softness", ctypes.c_double),
        ("size", ctypes.c_double * 3), ("h", ctypes.c_double), ("r", ctypes.c_double),
        ("ph", ctypes.c_double * 2), ("prot", ctypes.c_double * 4),
        ("lo", ctypes.c_double * 3), ("hi", ctypes.c_double * 3),
        ("min_gap", ctypes.c_double), ("max_gap", ctypes.c_double),
        ("action_scale", ctypes.c_double * 7),
    ]


class OrcColl(ctypes.Structure):
    _fields_ = [
        ("n_pairs", ctypes.c_int), ("n_points", ctypes.c_int),
        ("pairs", ctypes.c_void_p), ("rand", ctypes.c_void_p),
        ("idx", ctypes.c_void_p), ("pt", ctypes.c_void_p), ("pre", ctypes.c_void_p),
    ]


def build():
    """Compile liborc64.so / liborc32.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_LIBS = {}


def lib(bits=64):
    if bits not in _LIBS:
        path = os.path.join(HERE, "liborc%d.so" % bits)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        _LIBS[bits] = L
    return _LIBS[bits]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_sim(d):
    s = OrcSim()
    s.n_grid = int(d["n_grid"])
    s.n_prims = int(d.get("n_prims", 0))
    for k in ("dx", "inv_dx", "dt", "p_vol", "p_mass", "mu", "lam", "yield_stress",
              "ground_friction", "lower_bound"):
        setattr(s, k, float(d[k]))
    for i in range(3):
        s.gravity[i] = float(d["gravity"][i])
    return s


def make_prims(plist):
    arr = (OrcPrim * max(1, len(plist)))()
    for j, d in enumerate(plist):
        p = arr[j]
        p.type = int(d["type"]) if not isinstance(d["type"], str) else PRIM_TYPES[d["type"]]
        p.action_dim = int(d.get("action_dim", 0))
        p.friction = float(d.get("friction", 0.9))
        p.softness = float(d.get("softness", 666.0))
        for i, v in enumerate(d.get("size", (0.1, 0.1, 0.1))):
            p.size[i] = float(v)
        p.h = float(d.get("h", 0.0))
        p.r = float(d.get("r", 0.0))
        for i, v in enumerate(d.get("ph", (0.1, 0.1))):
            p.ph[i] = float(v)
        for i, v in enumerate(d.get("prot", (1.0, 0.0, 0.0, 0.0))):
            p.prot[i] = float(v)
        for i, v in enumerate(d.get("lo", (0.0, 0.0, 0.0))):
            p.lo[i] = float(v)
        for i, v in enumerate(d.get("hi", (1.0, 1.0, 1.0))):
            p.hi[i] = float(v)
        p.min_gap = float(d.get("min_gap", 0.0))
        p.max_gap = float(d.get("max_gap", 1.0))
        for i, v in enumerate(d.get("action_scale", ())):
            p.action_scale[i] = float(v)
    return arr


# per-class primitive defaults (Primitive.default_config subclasses, primitives.py)
PRIM_DEFAULTS = {
    "Sphere": dict(radius=1.0),
    "Capsule": dict(h=0.06, r=0.03),
    "RollingPin": dict(h=0.06, r=0.03),
    "RollingPinExt": dict(h=0.06, r=0.03),
    "Box": dict(size=(0.1, 0.1, 0.1)),
    "Gripper": dict(size=(0.03, 0.06, 0.03), minimal_gap=0.06, maximal_gap=1.0, init_gap=0.06),
    "Gripper2": dict(h=0.06, r=0.015, minimal_gap=0.06, maximal_gap=1.0, init_gap=0.06),
    "Chopsticks": dict(h=0.06, r=0.03, minimal_gap=0.06, init_gap=0.06),
    "Prism": dict(h=(0.1, 0.1), prot=(1.0, 0.0, 0.0, 0.0)),
    "Knife": dict(h=(0.1, 0.1), size=(0.1, 0.1, 0.1), prot=(1.0, 0.0, 0.0, 0.0)),
    "Cylinder": dict(h=0.2, r=0.1),
    "Torus": dict(tx=0.2, ty=0.1),
}


def sim_from_cfg(cfg):
    """SIMULATOR cfg -> oracle constants, the arithmetic of MPMSimulator.__init__
    (plb/engine/mpm_simulator.py:19-39) with the defaults of
    plb/config/default_config.py:11-22; optional overrides n_grid / dt / substeps."""
    d = dict(quality=1, quality_multiplier=1.0, yield_stress=50.0, E=5e3, nu=0.15, ground_friction=1.5,
             gravity=(0, -1, 0), lower_bound=0.0)
    d.update(cfg)
    q = d["quality"] * d["quality_multiplier"] * 0.5
    n_grid = int(d.get("n_grid") or int(128 * q))
    dx = 1 / n_grid
    dt = float(d.get("dt") or 0.5e-4 / q)
    p_vol = (dx * 0.5) ** 2
    E, nu = d["E"], d["nu"]
    sim = dict(n_grid=n_grid, dx=dx, inv_dx=float(n_grid), dt=dt, p_vol=p_vol, p_mass=p_vol,
               mu=E / (2 * (1 + nu)), lam=E * nu / ((1 + nu) * (1 - 2 * nu)), yield_stress=float(d["yield_stress"]),
               gravity=tuple(float(g) for g in d["gravity"]), ground_friction=float(d["ground_friction"]),
               lower_bound=float(d["lower_bound"]))
    substeps = int(d.get("substeps") or int(2e-3 // dt))
    return sim, substeps


def prim_from_cfg(cfg, softness=666.0):
    """PRIMITIVES entry -> oracle primitive parameters."""
    shape = cfg["shape"]
    d = dict(PRIM_DEFAULTS.get(shape, {}))
    d.update(cfg)
    act = cfg.get("action", {}) or {}
    o = dict(type=shape, action_dim=int(act.get("dim", 0)), friction=float(cfg.get("friction", 0.9)),
             softness=softness, lo=tuple(cfg.get("lower_bound", (0, 0, 0))), hi=tuple(cfg.get("upper_bound", (1, 1, 1))),
             action_scale=tuple(act.get("scale", ())))
    if shape == "Sphere":
        o["r"] = d["radius"]
    if shape in ("Capsule", "RollingPin", "RollingPinExt", "Gripper2", "Chopsticks", "Cylinder"):
        o["h"], o["r"] = d["h"], d["r"]
    if shape == "Torus":
        o["h"], o["r"] = d["tx"], d["ty"]
    if shape in ("Box", "Gripper", "Knife"):
        o["size"] = tuple(d["size"])
    if shape in ("Prism", "Knife"):
        o["ph"], o["prot"] = tuple(d["h"]), tuple(d["prot"])
    if shape in ("Gripper", "Gripper2", "Chopsticks"):
        o["min_gap"] = d["minimal_gap"]
        o["max_gap"] = d.get("maximal_gap", 1.0)
    return o


def prim_state_from_cfg(cfg):
    """Initial primitive frame state: init_pos + init_rot (+ init_gap)."""
    shape = cfg["shape"]
    d = dict(PRIM_DEFAULTS.get(shape, {}))
    d.update(cfg)
    s = np.zeros(8)
    s[:3] = cfg.get("init_pos", (0.3, 0.3, 0.3))
    s[3:7] = cfg.get("init_rot", (1.0, 0.0, 0.0, 0.0))
    if shape in ("Gripper", "Gripper2", "Chopsticks"):
        s[7] = d["init_gap"]
    return s


def reset_state(x):
    """reset_kernel: v = 0, F = I, C = 0 (mpm_simulator.py:402-410)."""
    x = np.asarray(x, np.float64)
    st = np.zeros((x.shape[0], 24))
    st[:, 0:3] = x
    st[:, 15:24] = np.eye(3).reshape(-1)
    return st


class Oracle:
    """One scene advanced and differentiated by the C restatement."""

    def __init__(self, sim, prims, bits=64, threads=1, collisions=None):
        """collisions: optional (pairs, rand) -- tool-tool collision pairs
        (i, j) in the reference's order and rand (n_pairs, n_points, 3)."""
        self.bits = bits
        self.dtype = np.float64 if bits == 64 else np.float32
        self.L = lib(bits)
        self.L.orc_set_threads(int(threads))
        sim = dict(sim)
        sim["n_prims"] = len(prims)
        self.sim_d = sim
        self.prims_d = list(prims)
        self.S = make_sim(sim)
        self.P = make_prims(self.prims_d)
        n = int(sim["n_grid"])
        G = n * n * n
        dt = self.dtype
        self.grid = np.zeros(G * 4, dt)
        self.gout = np.zeros(G * 3, dt)
        self.ggout = np.zeros(G * 3, dt)
        self.ggrid = np.zeros(G * 4, dt)
        self.np_ = len(prims)
        self.coll = None
        self.last_coll = None
        if collisions is not None and len(collisions[0]) > 0:
            pairs = np.ascontiguousarray(np.asarray(collisions[0], np.int32).reshape(-1, 2))
            rand = self.arr(collisions[1])
            assert rand.ndim == 3 and rand.shape[0] == pairs.shape[0] and rand.shape[2] == 3
            self.coll = (pairs, rand)

    def _coll(self, rec=None):
        """OrcColl over fresh (or the given) per-substep records."""
        pairs, rand = self.coll
        n = pairs.shape[0]
        if rec is None:
            rec = dict(idx=np.full(n, -1, np.int32), pt=np.zeros((n, 3), self.dtype), pre=np.zeros((n, 3), self.dtype))
        c = OrcColl()
        c.n_pairs = n
        c.n_points = rand.shape[1]
        c.pairs = pairs.ctypes.data
        c.rand = rand.ctypes.data
        c.idx = rec["idx"].ctypes.data
        c.pt = rec["pt"].ctypes.data
        c.pre = rec["pre"].ctypes.data
        return c, rec

    def tool_collide(self, ps1):
        """The collision pass alone on one frame's primitive states: returns
        (projected states, record dict idx / pt / pre)."""
        k = max(1, self.np_)
        ps1 = self.arr(ps1, (k, 8)).copy()
        c, rec = self._coll()
        self.L.orc_tool_collide(self.P, ctypes.byref(c), _ptr(ps1))
        return ps1, rec

    def tool_collide_rev(self, ps1, rec, gps1):
        k = max(1, self.np_)
        ps1 = self.arr(ps1, (k, 8))
        gps1 = self.arr(gps1, (k, 8)).copy()
        c, _ = self._coll(rec)
        self.L.orc_tool_collide_rev(self.P, ctypes.byref(c), _ptr(ps1), _ptr(gps1))
        return gps1

    def arr(self, a, shape=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if shape is not None:
            a = a.reshape(shape)
        return a

    # ---------------- single-step primitives ----------------
    def substep(self, st, ps_f, pv_f):
        st = self.arr(st)
        out = np.zeros_like(st)
        ps_f = self.arr(ps_f, (max(1, self.np_), 8))
        pv_f = self.arr(pv_f, (max(1, self.np_), 7))
        ps_f1 = np.zeros_like(ps_f)
        if self.coll is None:
            self.L.orc_substep(ctypes.byref(self.S), self.P, ctypes.c_int(st.shape[0]), _ptr(st), _ptr(out),
                               _ptr(ps_f), _ptr(pv_f), _ptr(ps_f1), _ptr(self.grid), _ptr(self.gout))
            self.last_coll = None
        else:
            c, rec = self._coll()
            self.L.orc_substep_c(ctypes.byref(self.S), self.P, ctypes.c_int(st.shape[0]), _ptr(st), _ptr(out),
                                 _ptr(ps_f), _ptr(pv_f), _ptr(ps_f1), _ptr(self.grid), _ptr(self.gout), ctypes.byref(c))
            self.last_coll = rec
        return out, ps_f1

    def substep_grad(self, st, adj1, ps_f, pv_f, ps_f1, gps_f1, coll=None):
        """Returns (adj0, gps_f contribution, gps_f1 (updated), gpv_f).
        coll: the collision record of this substep's forward (last_coll)."""
        st = self.arr(st)
        adj1 = self.arr(adj1)
        adj0 = np.zeros_like(st)
        k = max(1, self.np_)
        ps_f = self.arr(ps_f, (k, 8))
        pv_f = self.arr(pv_f, (k, 7))
        ps_f1 = self.arr(ps_f1, (k, 8))
        gps_f1 = self.arr(gps_f1, (k, 8)).copy()
        gps_f = np.zeros((k, 8), self.dtype)
        gpv_f = np.zeros((k, 7), self.dtype)
        args = [ctypes.byref(self.S), self.P, ctypes.c_int(st.shape[0]), _ptr(st), _ptr(adj1), _ptr(adj0), _ptr(ps_f),
                _ptr(pv_f), _ptr(ps_f1), _ptr(gps_f), _ptr(gps_f1), _ptr(gpv_f), _ptr(self.grid), _ptr(self.gout),
                _ptr(self.ggout), _ptr(self.ggrid)]
        if self.coll is None:
            self.L.orc_substep_grad(*args)
        else:
            assert coll is not None, "substep_grad of a colliding scene needs the forward's collision record"
            c, _ = self._coll(coll)
            self.L.orc_substep_grad_c(*args, ctypes.byref(c))
        return adj0, gps_f, gps_f1, gpv_f

    def grid_mass(self, st):
        st = self.arr(st)
        n = self.S.n_grid
        m = np.zeros(n * n * n, self.dtype)
        self.L.orc_grid_mass(ctypes.byref(self.S), ctypes.c_int(st.shape[0]), _ptr(st), _ptr(m))
        return m.reshape(n, n, n)

    def grid_mass_grad(self, st, gm):
        st = self.arr(st)
        gm = self.arr(gm).reshape(-1)
        adj = np.zeros_like(st)
        self.L.orc_grid_mass_grad(ctypes.byref(self.S), ctypes.c_int(st.shape[0]), _ptr(st), _ptr(gm), _ptr(adj))
        return adj

    def fk(self, ps_f, pv_f):
        k = max(1, self.np_)
        ps_f = self.arr(ps_f, (k, 8))
        pv_f = self.arr(pv_f, (k, 7))
        out = np.zeros_like(ps_f)
        self.L.orc_fk(self.P, ctypes.c_int(self.np_), _ptr(ps_f), _ptr(pv_f), _ptr(out))
        return out

    def fk_rev(self, ps_f, pv_f, gps_f1):
        k = max(1, self.np_)
        ps_f = self.arr(ps_f, (k, 8))
        pv_f = self.arr(pv_f, (k, 7))
        gps_f1 = self.arr(gps_f1, (k, 8))
        gps_f = np.zeros((k, 8), self.dtype)
        gpv_f = np.zeros((k, 7), self.dtype)
        self.L.orc_fk_rev(self.P, ctypes.c_int(self.np_), _ptr(ps_f), _ptr(pv_f), _ptr(gps_f1), _ptr(gps_f), _ptr(gpv_f))
        return gps_f, gpv_f

    def set_velocity(self, j, a, n_substeps):
        pv = np.zeros(7, self.dtype)
        a = self.arr(np.asarray(a).reshape(-1))
        if a.size == 0:
            return pv
        self.L.orc_set_velocity(ctypes.byref(self.P[j]), _ptr(a), ctypes.c_int(n_substeps), _ptr(pv))
        return pv

    def min_dist(self, st, ps):
        """compute_min_dist: (n_cols, n) distances of every particle to every
        primitive (both jaws of a two-jaw gripper)."""
        st = self.arr(st)
        k = max(1, self.np_)
        ps = self.arr(ps, (k, 8))
        nc = self.L.orc_min_dist_cols(self.P, ctypes.c_int(self.np_))
        out = np.zeros((nc, st.shape[0]), self.dtype)
        self.L.orc_min_dist(self.P, ctypes.c_int(self.np_), ctypes.c_int(st.shape[0]), _ptr(st), _ptr(ps), _ptr(out))
        return out

    def min_dist_grad(self, st, ps, gout):
        """Adjoint of min_dist: (particle adjoint (n, 24) with the x rows set,
        primitive-state adjoint (n_prims, 8))."""
        st = self.arr(st)
        k = max(1, self.np_)
        ps = self.arr(ps, (k, 8))
        gout = self.arr(gout)
        adj = np.zeros_like(st)
        gps = np.zeros((k, 8), self.dtype)
        self.L.orc_min_dist_grad(self.P, ctypes.c_int(self.np_), ctypes.c_int(st.shape[0]), _ptr(st), _ptr(ps), _ptr(gout),
                                 _ptr(adj), _ptr(gps))
        return adj, gps

    def prim_sdf(self, j, ps, x):
        x = self.arr(x, (-1, 3))
        ps = self.arr(ps, (8,))
        out = np.zeros(x.shape[0], self.dtype)
        self.L.orc_prim_sdf(ctypes.byref(self.P[j]), _ptr(ps), ctypes.c_int(x.shape[0]), _ptr(x), _ptr(out))
        return out

    def prim_normal(self, j, ps, x):
        x = self.arr(x, (-1, 3))
        ps = self.arr(ps, (8,))
        out = np.zeros_like(x)
        self.L.orc_prim_normal(ctypes.byref(self.P[j]), _ptr(ps), ctypes.c_int(x.shape[0]), _ptr(x), _ptr(out))
        return out

    def grid_node(self, ps0, ps1, I, pm):
        k = max(1, self.np_)
        ps0 = self.arr(ps0, (k, 8))
        ps1 = self.arr(ps1, (k, 8))
        I = np.ascontiguousarray(I, dtype=np.int32)
        pm = self.arr(pm, (4,))
        out = np.zeros(3, self.dtype)
        self.L.orc_grid_node(ctypes.byref(self.S), self.P, _ptr(ps0), _ptr(ps1), _ptr(I), _ptr(pm), _ptr(out))
        return out

    def grid_node_rev(self, ps0, ps1, I, pm, gvout):
        k = max(1, self.np_)
        ps0 = self.arr(ps0, (k, 8))
        ps1 = self.arr(ps1, (k, 8))
        I = np.ascontiguousarray(I, dtype=np.int32)
        pm = self.arr(pm, (4,))
        gvout = self.arr(gvout, (3,))
        gpm = np.zeros(4, self.dtype)
        g0 = np.zeros((k, 8), self.dtype)
        g1 = np.zeros((k, 8), self.dtype)
        self.L.orc_grid_node_rev(ctypes.byref(self.S), self.P, _ptr(ps0), _ptr(ps1), _ptr(I), _ptr(pm),
                                 _ptr(gvout), _ptr(gpm), _ptr(g0), _ptr(g1))
        return gpm, g0, g1

    def svd(self, A):
        A = self.arr(A, (3, 3))
        U = np.zeros((3, 3), self.dtype)
        S = np.zeros((3, 3), self.dtype)
        V = np.zeros((3, 3), self.dtype)
        self.L.orc_svd3(_ptr(A), _ptr(U), _ptr(S), _ptr(V))
        return U, S, V

    def backward_svd(self, gu, gs, gv, u, s, v):
        args = [self.arr(a, (3, 3)) for a in (gu, gs, gv, u, s, v)]
        out = np.zeros((3, 3), self.dtype)
        self.L.orc_backward_svd(*[_ptr(a) for a in args], _ptr(out))
        return out

    def particle_stress(self, st):
        st = self.arr(st)
        Fn = np.zeros((st.shape[0], 9), self.dtype)
        stress = np.zeros((st.shape[0], 9), self.dtype)
        self.L.orc_particle_stress(ctypes.byref(self.S), ctypes.c_int(st.shape[0]), _ptr(st), _ptr(Fn), _ptr(stress))
        return Fn, stress


class OracleRollout:
    """Store-all trajectory driver mirroring ``GradModel.forward_step`` /
    ``backward_step`` (plb/engine/function.py:166-190): actions are set per
    env step (clipped to [-1, 1]), ``substeps`` substeps run per env step."""

    def __init__(self, oracle, substeps):
        self.o = oracle
        self.substeps = substeps

    def run(self, st0, ps0, actions):
        o = self.o
        k = max(1, o.np_)
        H = len(actions)
        S = self.substeps
        self.actions = [np.clip(np.asarray(a, np.float64).reshape(-1), -1, 1) for a in actions]
        self.frames = [o.arr(st0)]
        self.ps = [o.arr(ps0, (k, 8))]
        self.pv = []
        self.colls = []
        dims = [int(p.get("action_dim", 0)) for p in o.prims_d]
        offs = np.concatenate([[0], np.cumsum(dims)]).astype(int)
        for s in range(H):
            pv_step = np.zeros((k, 7), o.dtype)
            for j in range(o.np_):
                if dims[j] > 0:
                    pv_step[j] = o.set_velocity(j, self.actions[s][offs[j]:offs[j + 1]], S)
            for i in range(S):
                self.pv.append(pv_step.copy())
                st1, ps1 = o.substep(self.frames[-1], self.ps[-1], pv_step)
                self.colls.append(o.last_coll)
                self.frames.append(st1)
                self.ps.append(ps1)
        self.offs = offs
        self.dims = dims
        return self.frames, self.ps

    def grad(self, frame_adj, prim_adj=None):
        """frame_adj: dict frame -> (n, 24) adjoint injected at that frame
        (like GradModel.set_obs_grad).  Returns (d/d actions (H, A), adj of
        frame 0, primitive-state adjoint of frame 0)."""
        o = self.o
        k = max(1, o.np_)
        T = len(self.frames) - 1
        S = self.substeps
        H = T // S
        adj = np.zeros_like(self.frames[0])
        gps = [np.zeros((k, 8), o.dtype) for _ in range(T + 1)]
        if prim_adj:
            for f, g in prim_adj.items():
                gps[f] += g
        gpv = [np.zeros((k, 7), o.dtype) for _ in range(T)]
        if T in frame_adj:
            adj = adj + frame_adj[T]
        for f in range(T - 1, -1, -1):
            adj0, g0, g1, gv = o.substep_grad(self.frames[f], adj, self.ps[f], self.pv[f], self.ps[f + 1], gps[f + 1],
                                              self.colls[f])
            gps[f] += g0
            gps[f + 1] = g1
            gpv[f] = gv
            if f in frame_adj:
                adj0 = adj0 + frame_adj[f]
            adj = adj0
        A = int(sum(self.dims))
        ga = np.zeros((H, A), np.float64)
        for s in range(H):
            for j in range(o.np_):
                d = self.dims[j]
                if d == 0:
                    continue
                sc = np.zeros(7)
                sc[:len(o.prims_d[j]["action_scale"])] = o.prims_d[j]["action_scale"]
                acc = np.zeros(7)
                for i in range(s * S, (s + 1) * S):
                    acc += gpv[i][j]
                g = np.zeros(d)
                g[:min(d, 3)] = (acc[:3] * sc[:3] / S)[:min(d, 3)]
                if d > 3:
                    g[3:6] = acc[3:6] * sc[3:6] / S
                if d > 6:
                    g[6] = acc[6] * sc[6] / S
                ga[s, self.offs[j]:self.offs[j + 1]] = g
        return ga, adj, gps[0]
