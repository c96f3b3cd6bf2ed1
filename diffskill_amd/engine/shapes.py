"""Initial particle clouds from a task's SHAPES list: the Shapes class of
plb/engine/shapes/shape_maker.py:13-168.

Every shape draws from numpy's global legacy generator with the reference's
sequence of calls (``scatter`` reseeds it), and the generator state is put
back when the constructor ends, so the same global state yields the same
particles as the reference.  Particle counts follow get_n_particles (30000
per 0.1^3 of volume, at least 1); ``init_rot`` rotates an object about its
mean with the rotation matrix of a (w, x, y, z) quaternion.
"""
import ast

import numpy as np

COLORS = [(127 << 16) + 127, (127 << 8), 127, 127 << 16]


def quat2mat(q):
    """Rotation matrix of the quaternion (w, x, y, z), normalised implicitly
    (the published transforms3d.quaternions.quat2mat formula)."""
    w, x, y, z = [float(v) for v in q]
    n = w * w + x * x + y * y + z * z
    if n < np.finfo(np.float64).eps:
        return np.eye(3)
    s = 2.0 / n
    X, Y, Z = x * s, y * s, z * s
    wX, wY, wZ = w * X, w * Y, w * Z
    xX, xY, xZ = x * X, x * Y, x * Z
    yY, yZ, zZ = y * Y, y * Z, z * Z
    return np.array([[1.0 - (yY + zZ), xY - wZ, xZ + wY],
                     [xY + wZ, 1.0 - (xX + zZ), yZ - wX],
                     [xZ - wY, yZ + wX, 1.0 - (xX + yY)]])


_BINOPS = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b, ast.Mult: lambda a, b: a * b,
           ast.Div: lambda a, b: a / b, ast.FloorDiv: lambda a, b: a // b, ast.Mod: lambda a, b: a % b,
           ast.Pow: lambda a, b: a ** b, ast.LShift: lambda a, b: a << b, ast.RShift: lambda a, b: a >> b,
           ast.BitOr: lambda a, b: a | b, ast.BitAnd: lambda a, b: a & b, ast.BitXor: lambda a, b: a ^ b}


def _eval_node(n):
    if isinstance(n, ast.Expression):
        return _eval_node(n.body)
    if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
        return n.value
    if isinstance(n, (ast.Tuple, ast.List)):
        vals = [_eval_node(e) for e in n.elts]
        return tuple(vals) if isinstance(n, ast.Tuple) else vals
    if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
        v = _eval_node(n.operand)
        return -v if isinstance(n.op, ast.USub) else v
    if isinstance(n, ast.BinOp) and type(n.op) in _BINOPS:
        return _BINOPS[type(n.op)](_eval_node(n.left), _eval_node(n.right))
    raise ValueError("unsupported expression in a SHAPES value: %s" % ast.dump(n))


def _value(v):
    """A SHAPES value; strings are the Python expressions the reference
    evaluates (shape_maker.py:24), e.g. "(0.5, 0.1, 0.5)" or
    "((150<<8)+ (150<<16))": numbers, tuples, lists and arithmetic / bit
    operators only, evaluated without running code."""
    return _eval_node(ast.parse(v.strip(), mode="eval")) if isinstance(v, str) else v


class Shapes:
    def __init__(self, cfg):
        self.objects = []
        self.colors = []
        self.dim = 3
        state = np.random.get_state()
        try:
            for item in cfg:
                kind = item["shape"]
                kw = {k: _value(v) for k, v in item.items() if k != "shape"}
                if kind == "box":
                    self.add_box(**kw)
                elif kind == "multibox":
                    self.add_multibox(**kw)
                elif kind == "sphere":
                    self.add_sphere(**kw)
                elif kind == "multisphere":
                    self.add_multisphere(**kw)
                elif kind == "capsule":
                    self.add_capsule(**kw)
                elif kind == "cylinder":
                    self.add_cylinder(**kw)
                elif kind == "scatter":
                    self.add_scatter(**kw)
                else:
                    raise NotImplementedError("Shape %s is not supported!" % kind)
        finally:
            np.random.set_state(state)

    def get_n_particles(self, volume):
        return max(int(volume / (0.1 ** 3) * 30000), 1)

    def add_object(self, particles, color=None, init_rot=None):
        particles = np.asarray(particles, np.float64)
        if init_rot is not None:
            origin = particles.mean(axis=0)
            particles = (particles[:, :self.dim] - origin) @ quat2mat(init_rot).T + origin
        self.objects.append(particles[:, :self.dim])
        if color is None or isinstance(color, int):
            c = np.zeros(len(particles), np.int32)
            c[:] = COLORS[len(self.objects) - 1] if color is None else color
            color = c
        self.colors.append(color)

    def add_box(self, init_pos, width, n_particles=None, color=None, init_rot=None):
        width = np.array([width] * self.dim) if isinstance(width, float) else np.array(width)
        if n_particles is None:
            n_particles = self.get_n_particles(np.prod(width))
        p = (np.random.random((n_particles, self.dim)) * 2 - 1) * (0.5 * width) + np.array(init_pos)
        self.add_object(p, color, init_rot=init_rot)

    def add_multibox(self, all_pos, all_width, all_rot, color):
        volumes = [np.prod(w) for w in all_width]
        total = sum(volumes)
        n = min(30000, self.get_n_particles(total))
        for pos, width, rot, vol in zip(all_pos, all_width, all_rot, volumes):
            self.add_box(pos, width, n_particles=int(n * vol / total), color=color, init_rot=rot)

    def add_sphere(self, init_pos, radius, n_particles=None, color=None, init_rot=None):
        if n_particles is None:
            n_particles = self.get_n_particles((radius ** 3) * 4 * np.pi / 3)
        p = np.random.normal(size=(n_particles, self.dim))
        p /= np.linalg.norm(p, axis=-1, keepdims=True)
        u = np.random.random(size=(n_particles, 1)) ** (1. / self.dim)
        p = p * u * radius + np.array(init_pos)[:self.dim]
        self.add_object(p, color, init_rot=init_rot)

    def add_multisphere(self, all_pos, all_r, color):
        volumes = [(r ** 3) * 4 * np.pi / 3 for r in all_r]
        total = sum(volumes)
        n = self.get_n_particles(total)
        for pos, r, vol in zip(all_pos, all_r, volumes):
            self.add_sphere(pos, r, n_particles=int(n * vol / total))

    def add_capsule(self, init_pos, radius, height, n_particles=None, color=None, init_rot=None):
        # the default count is that of the sphere alone (shape_maker.py:114-119)
        if n_particles is None:
            n_particles = self.get_n_particles((radius ** 3) * 4 * np.pi / 3)
        v_sphere = (radius ** 3) * 4 * np.pi / 3
        v_cyl = (radius ** 2) * np.pi * height
        n1 = int(n_particles * v_sphere / (v_sphere + v_cyl))
        n2 = n_particles - n1
        p = np.random.normal(size=(n1, self.dim))
        p /= np.linalg.norm(p, axis=-1, keepdims=True)
        u = np.random.random(size=(n1, 1)) ** (1. / self.dim)
        p = p * u * radius
        p[:, 2] += np.sign(p[:, 2]) * height / 2.
        p += np.array(init_pos)[:self.dim]
        p = np.vstack([p, self._disk_column(radius, height, n2, init_pos)])
        self.add_object(p, color, init_rot=init_rot)

    def add_cylinder(self, init_pos, radius, height, n_particles=None, color=None, init_rot=None):
        if n_particles is None:
            n_particles = self.get_n_particles((radius ** 3) * 4 * np.pi / 3)
        self.add_object(self._disk_column(radius, height, n_particles, init_pos), color, init_rot=init_rot)

    @staticmethod
    def _disk_column(radius, height, n, init_pos):
        # uniform in a disk of the given radius (x, y), uniform along z
        r = radius * np.sqrt(np.random.random([n, 1]))
        theta = np.random.random([n, 1]) * 2 * np.pi
        x, y = (np.cos(theta) * r).reshape(-1, 1), (np.sin(theta) * r).reshape(-1, 1)
        h = np.random.random([n, 1]) * height - height / 2.
        return np.hstack([x, y, h]) + init_pos

    def add_scatter(self, seed, pos_min, pos_max, color=None, **_):
        """40 seeded centres in [pos_min, pos_max], 50 particles each with
        Gaussian noise (0.004; height 0.2x and +0.1), shape_maker.py:37-50."""
        np.random.seed(seed)
        N, multiply = 40, 50
        centres = np.hstack([np.random.uniform(lo, hi, size=N).reshape([N, 1]) for lo, hi in zip(pos_min, pos_max)])
        noise = np.random.normal(0, scale=0.004, size=N * 3 * multiply).reshape(multiply, N, 3)
        noise[:, :, 1] = noise[:, :, 1] * 0.2 + 0.1
        self.add_object((centres.reshape([1, N, 3]) + noise).reshape(multiply * N, 3), color=color)

    def get(self):
        assert len(self.objects) > 0, "please add at least one shape into the scene"
        return np.concatenate(self.objects), np.concatenate(self.colors)
