"""Configuration: the reference's SIMULATOR / PRIMITIVES keys and the
constants MPMSimulator derives from them.

The reference uses yacs ``CfgNode`` (plb/config/default_config.py,
plb/config/utils.py); yacs is not a dependency here, so ``CfgNode`` below is
a small attribute-dict with the same merge behaviour for the keys the
simulator reads (strings that spell Python literals, e.g. YAML
``(0.5, 0.05, 0.5)``, are decoded like yacs' ``_decode_cfg_value``).
"""
import ast
import copy


class CfgNode(dict):
    """Attribute-access dict (the subset of yacs.CfgNode the engine uses)."""

    def __init__(self, init=None, **kw):
        super().__init__()
        for k, v in dict(init or {}, **kw).items():
            self[k] = _wrap(v)

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = _wrap(v)

    def clone(self):
        return copy.deepcopy(self)

    def merge(self, other):
        """Recursive merge of a dict / CfgNode (new keys allowed)."""
        for k, v in (other or {}).items():
            v = _decode(v)
            if isinstance(v, dict) and isinstance(self.get(k), CfgNode):
                self[k].merge(v)
            else:
                self[k] = _wrap(v)
        return self

    merge_from_other_cfg = merge

    def merge_from_file(self, path):
        import yaml
        with open(path) as f:
            return self.merge(yaml.safe_load(f))

    def merge_from_list(self, opts):
        for k, v in zip(opts[0::2], opts[1::2]):
            node = self
            parts = k.split(".")
            for p in parts[:-1]:
                node = node[p]
            node[parts[-1]] = _wrap(_decode(v))
        return self

    def defrost(self):
        return self

    def freeze(self):
        return self


def _decode(v):
    if isinstance(v, str):
        try:
            return ast.literal_eval(v)
        except (ValueError, SyntaxError):
            return v
    return v


def _wrap(v):
    v = _decode(v)
    if isinstance(v, CfgNode):
        return v
    if isinstance(v, dict):
        return CfgNode(v)
    if isinstance(v, list):
        return [(_wrap(i) if isinstance(i, dict) else _decode(i)) for i in v]
    return v


def simulator_defaults():
    """SIMULATOR defaults (plb/config/default_config.py:11-22)."""
    return CfgNode(dict(
        dim=3, quality=1, quality_multiplier=1., yield_stress=50., dtype="float32",
        max_steps=1024, n_particles=30000, lower_bound=0., E=5e3, nu=0.15,
        ground_friction=1.5, gravity=(0, -1, 0),
    ))


def get_cfg_defaults():
    cfg = CfgNode()
    cfg.SIMULATOR = simulator_defaults()
    cfg.PRIMITIVES = []
    cfg.SHAPES = []
    cfg.ENV = CfgNode(dict(loss=dict(soft_contact=False, weight=dict(sdf=10, density=10, contact=1), target_path=""),
                           n_observed_particles=200, cached_state_path="", env_name=""))
    return cfg


def load(path=None, opts=None):
    """plb.config.load (plb/config/utils.py:33-41)."""
    cfg = get_cfg_defaults()
    if path is not None:
        cfg.merge_from_file(path)
    if opts is not None:
        cfg.merge_from_list(opts)
    return cfg


def derive(sim_cfg):
    """Constants of MPMSimulator.__init__ (plb/engine/mpm_simulator.py:19-39),
    computed in double precision like the reference's Python.

    Extra keys accepted (absent in the reference, used by workloads that
    override the derived values): ``n_grid``, ``dt``, ``substeps``.
    """
    c = simulator_defaults()
    c.merge(sim_cfg)
    dim = int(c.dim)
    if dim != 3:
        raise NotImplementedError("only the 3D simulator is implemented (dim=3)")
    quality = c.quality * c.quality_multiplier
    quality = quality * 0.5
    n_grid = int(c.get("n_grid", None) or int(128 * quality))
    dx, inv_dx = 1 / n_grid, float(n_grid)
    dt = float(c.get("dt", None) or (0.5e-4 / quality))
    p_vol = (dx * 0.5) ** 2
    p_rho = 1
    E, nu = c.E, c.nu
    out = dict(
        dim=dim, n_grid=n_grid, dx=dx, inv_dx=inv_dx, dt=dt, p_vol=p_vol, p_rho=p_rho, p_mass=p_vol * p_rho,
        mu=E / (2 * (1 + nu)), lam=E * nu / ((1 + nu) * (1 - 2 * nu)),
        yield_stress=float(c.yield_stress), ground_friction=float(c.ground_friction),
        gravity=tuple(float(g) for g in c.gravity), lower_bound=float(c.lower_bound),
        max_steps=int(c.max_steps), n_particles=int(c.n_particles), dtype=c.dtype,
    )
    out["substeps"] = int(c.get("substeps", None) or int(2e-3 // dt))
    return out
