"""Host-side logic that runs without a GPU: the C ABI library loads and exports
every symbol the public header declares; cfg handling and constant derivation
of the drop-in API; primitive construction from reference-style cfgs."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    syms = set()
    for name in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if name.endswith(".h"):
            src = open(os.path.join(ROOT, "include", name)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            syms |= set(re.findall(r"\b((?:mpm|pc)_[a-z0-9_]+)\s*\(", src))
    return sorted(syms)


def test_library_exports_every_header_symbol():
    from diffskill_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.EXPORTS)


def test_binding_signatures_and_error_string():
    from diffskill_amd import _lib
    L = _lib.load()
    assert L.mpm_last_error() is not None
    # host-side validation only: no device work
    assert L.mpm_workspace_bytes(None, 0, 10, 0, -1) == -1
    d = _lib.SceneDesc()
    d.n_grid, d.n_particles, d.n_prims = 64, 1000, 1
    assert L.mpm_workspace_bytes(ctypes.byref(d), 1, 101, 0, -1) > 0
    store_all = L.mpm_workspace_bytes(ctypes.byref(d), 1, 1001, 0, -1)
    ckpt = L.mpm_workspace_bytes(ctypes.byref(d), 1, 1001, 50, -1)
    assert ckpt < store_all
    # kept grids and bins cost memory per frame slot; 0 keeps nothing
    assert L.mpm_workspace_bytes(ctypes.byref(d), 1, 1001, 0, 0) < store_all
    assert L.mpm_workspace_bytes(ctypes.byref(d), 1, 1001, 0, 64) < L.mpm_workspace_bytes(ctypes.byref(d), 1, 1001, 0, 128)
    rc = L.mpm_create(ctypes.byref(d), 1, None, 1, 0, -1, None, 0, None, ctypes.byref(ctypes.c_void_p()))
    assert rc != 0 and b"max_frames" in L.mpm_last_error()


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors have the sizes and field offsets the C compiler gives
    the header's structs."""
    import subprocess
    from diffskill_amd import _lib
    fields = {"mpm_scene_desc": [f[0] for f in _lib.SceneDesc._fields_], "mpm_prim_desc": [f[0] for f in _lib.PrimDesc._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "diffskill_mpm.h"', "int main(void) {"]
    for st, fs in fields.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (st, st))
        for f in fs:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (st, f, st, f))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for st, cls in (("mpm_scene_desc", _lib.SceneDesc), ("mpm_prim_desc", _lib.PrimDesc)):
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f in fields[st]:
            assert int(out["%s.%s" % (st, f)]) == getattr(cls, f).offset, (st, f)


def test_config_derivation_matches_reference_arithmetic():
    from diffskill_amd.config import derive
    c = derive({})
    assert c["n_grid"] == 64 and c["dt"] == 1e-4 and c["substeps"] == 19
    assert c["p_mass"] == (0.5 / 64) ** 2
    c = derive(dict(quality_multiplier=1.25))
    assert c["n_grid"] == 80 and c["substeps"] == 24
    c = derive(dict(dt=5e-4, substeps=10, n_grid=64))
    assert c["dt"] == 5e-4 and c["substeps"] == 10
    with pytest.raises(NotImplementedError):
        derive(dict(dim=2))


def test_cfgnode_decodes_yaml_literals():
    import yaml
    from diffskill_amd.config import get_cfg_defaults
    src = """
SIMULATOR:
  E: 5000.
  yield_stress: 50.
PRIMITIVES:
  - shape: RollingPin
    h: 0.3
    r: 0.03
    init_pos: (0.5, 0.123, 0.5)
    init_rot: (0.707, 0.707, 0., 0.)
    action:
      dim: 3
      scale: (0.6666666666666667, 0.06666666666666668, 0.001)
"""
    cfg = get_cfg_defaults()
    cfg.merge(yaml.safe_load(src))
    assert cfg.SIMULATOR.E == 5000.0 and cfg.SIMULATOR.nu == 0.15
    p = cfg.PRIMITIVES[0]
    assert p.init_pos == (0.5, 0.123, 0.5)
    assert p.action.scale[2] == 0.001


def test_primitives_container_api():
    from diffskill_amd.engine.primitive import Primitives
    prims = Primitives([
        dict(shape="RollingPin", h=0.3, r=0.03, init_pos=(0.5, 0.12, 0.5), action=dict(dim=3, scale=(0.6, 0.06, 0.001))),
        dict(shape="Gripper", size=(0.015, 0.09, 0.05), init_gap=0.4, minimal_gap=0.05,
             action=dict(dim=7, scale=(0.015, 0.0, 0.015, 0.0, 0.0, 0.1, 0.03))),
        dict(shape="Box", size=(0.07, 0.07, 0.02)),
    ])
    assert len(prims) == 3 and prims.action_dims == [0, 3, 10, 10] and prims.action_dim == 10
    assert prims.state_dims == [7, 8, 7] and prims.state_dim == 22
    np.testing.assert_allclose(prims[1].init_state, (0.3, 0.3, 0.3, 1, 0, 0, 0, 0.4))
    d = prims[1].desc()
    assert d.type == 5 and d.min_gap == 0.05 and d.max_gap == 1.0 and tuple(d.size) == (0.015, 0.09, 0.05)
    # state kept on the host until a simulator context exists
    prims[0].set_state(3, [0.1, 0.2, 0.3])
    np.testing.assert_allclose(prims[0].get_state(3)[:3], [0.1, 0.2, 0.3])
    cyl, tor = Primitives([dict(shape="Cylinder", h=0.05, r=0.02), dict(shape="Torus")])
    d = cyl.desc()
    assert d.type == 10 and (d.h, d.r) == (0.05, 0.02)
    d = tor.desc()
    assert d.type == 11 and (d.h, d.r) == (0.2, 0.1)     # Torus default tx, ty (primitives.py:352-356)
    with pytest.raises(NotImplementedError):
        Primitives([dict(shape="Cone")])


def test_collision_pairs_and_rand_num():
    """set_num_collision (mpm_simulator.py:69-80) on the PRIMITIVES of
    plb/envs/gather_move.yml (Gripper and Lifter Box with collision_group
    (0, 0, 1), Obstacle Box without) and lift_spread.yml (only the Box has
    one); rand_num of initialize (mpm_simulator.py:87-96) is what
    np.random.seed(42); np.random.uniform(-1.5, 1.5, ...) draws, and the
    caller's global random state is left as it was."""
    from diffskill_amd.engine.mpm_simulator import collision_pairs, collision_rand
    from diffskill_amd.engine.primitive import Primitives
    gather = Primitives([
        dict(shape="Gripper", size=(0.015, 0.09, 0.05), init_gap=0.4, minimal_gap=0.05, collision_group="(0, 0, 1)",
             action=dict(dim=7, scale=(0.015, 0.0, 0.015, 0.0, 0.0, 0.1, 0.03))),
        dict(shape="Box", size=(0.07, 0.07, 0.02), collision_group="(0, 0, 1)", action=dict(dim=6, scale=(0.01,) * 6)),
        dict(shape="Box", size=(0.2, 0.28, 0.04)),
    ])
    assert collision_pairs([p.cfg for p in gather]) == [(0, 2), (1, 2)]
    lift = Primitives([
        dict(shape="RollingPinExt", h=0.3, r=0.03, action=dict(dim=6, scale=(0.7, 0.005, 0.005, 0.005, 0., 0.))),
        dict(shape="Box", size=(0.1, 0.1, 0.02), collision_group=(0, 0, 1), action=dict(dim=6, scale=(0.01,) * 6)),
        dict(shape="Box", size=(0.2, 0.28, 0.07)),
    ])
    assert collision_pairs([p.cfg for p in lift]) == [(1, 2)]
    # five primitives with the three-entry default: missing entries count as 0
    five = [dict(shape="Sphere", radius=0.05)] * 4 + [dict(shape="Box", collision_group=(0, 0, 0, 1, 0))]
    assert collision_pairs([p.cfg for p in Primitives(five)]) == [(4, 3)]
    np.random.seed(123)
    before = np.random.get_state()[1].copy()
    r = collision_rand(2)
    assert np.array_equal(np.random.get_state()[1], before)
    np.random.seed(42)
    ref = np.random.uniform(-1.5, 1.5, (2, 600, 3))
    assert r.dtype == np.float32 and r.shape == (2, 600, 3)
    np.testing.assert_array_equal(r, ref.astype(np.float32))


def test_workloads_are_seeded_and_sized():
    from diffskill_amd import workloads
    w = workloads.config4()
    assert len(w["scenes"]) == 64 and w["H"] * w["substeps"] == 200
    assert w["sizes"].min() >= 2000 and w["sizes"].max() <= 40000
    w2 = workloads.config4()
    np.testing.assert_array_equal(w["sizes"], w2["sizes"])
    np.testing.assert_array_equal(w["scenes"][5]["x0"], w2["scenes"][5]["x0"])
    c2 = workloads.config2()
    assert c2["x0"].shape == (20000, 3) and c2["H"] * c2["substeps"] == 500


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built HIP library the product path raises; there is no
    CPU fallback to route through."""
    from diffskill_amd import _lib
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libdiffskill_mpm.so"))
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()
