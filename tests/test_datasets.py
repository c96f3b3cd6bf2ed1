"""DiffSkill dataset formats (state_generator.py:34-67, multitask_env.py:22-32,
70-91): lzma-pickled TaichiEnv state dicts, target point clouds and images,
read and written on the host."""
import lzma
import os
import pickle

import numpy as np
import pytest

from diffskill_amd import datasets


def _state(n=50, seed=0, gap=True):
    rng = np.random.default_rng(seed)
    prims = [rng.normal(size=7), rng.normal(size=8)] if gap else [rng.normal(size=7)]
    return {"state": [rng.uniform(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3, 3)),
                      rng.normal(size=(n, 3, 3))] + prims,
            "softness": 666.0, "is_copy": True}


def _same(a, b):
    assert a.keys() == b.keys() and a["softness"] == b["softness"] and a["is_copy"] == b["is_copy"]
    assert len(a["state"]) == len(b["state"])
    for u, v in zip(a["state"], b["state"]):
        assert u.dtype == v.dtype and u.shape == v.shape
        np.testing.assert_array_equal(u, v)


def test_state_file_round_trip(tmp_path):
    s = _state()
    p = str(tmp_path / "init" / "state_0.xz")
    datasets.save_state(p, s)
    with open(p, "rb") as f:
        assert f.read(6) == b"\xfd7zXZ\x00"            # an .xz container, as lzma.open writes
    _same(datasets.load_state(p), s)
    # what the reference's reader does with the same file
    with lzma.open(p, "rb") as f:
        _same(pickle.load(f), s)


def test_state_file_written_with_numpy1_module_names(tmp_path):
    """Files written under numpy 1.x name numpy.core.multiarray; they load."""
    s = _state(gap=False)
    raw = pickle.dumps(s, protocol=2)       # GLOBAL opcodes spell module and name as text
    assert b"numpy._core.multiarray" in raw or b"numpy.core.multiarray" in raw
    raw = raw.replace(b"numpy._core.multiarray", b"numpy.core.multiarray")
    p = str(tmp_path / "s.xz")
    with lzma.open(p, "wb") as f:
        f.write(raw)
    _same(datasets.load_state(p), s)


def test_state_file_with_foreign_globals_is_refused(tmp_path):
    """A pickle that names any global besides numpy's reconstructors is
    refused when the global is looked up, before anything is called."""
    marker = tmp_path / "ran"
    payload = b"cos\nsystem\n(S'touch " + str(marker).encode() + b"'\ntR."
    p = str(tmp_path / "evil.xz")
    with lzma.open(p, "wb") as f:
        f.write(payload)
    with pytest.raises(pickle.UnpicklingError, match="os.system"):
        datasets.load_state(p)
    assert not marker.exists()
    p2 = str(tmp_path / "notdict.xz")
    with lzma.open(p2, "wb") as f:
        pickle.dump([1, 2, 3], f, protocol=4)
    with pytest.raises(ValueError):
        datasets.load_state(p2)


def test_cached_state_layout_and_natural_order(tmp_path):
    root = str(tmp_path / "ds")
    cs = datasets.CachedStates(root)
    for i in (0, 2, 10, 1):
        cs.save_init(i, _state(seed=i))
        cs.save_target(i, np.full((20, 3), float(i)))
    np.save(os.path.join(root, "target", "target_imgs.npy"), np.zeros((4, 8, 8, 4), np.float32))
    cs = datasets.CachedStates(root)
    assert cs.num_inits == 4 and cs.num_targets == 4
    assert [os.path.basename(p) for p in cs.target_paths] == ["target_0.npy", "target_1.npy", "target_2.npy", "target_10.npy"]
    _same(cs.init_state(10), _state(seed=10))
    pcs = cs.target_pcs()
    assert pcs.shape == (4, 20, 3) and [pcs[k, 0, 0] for k in range(4)] == [0.0, 1.0, 2.0, 10.0]
    np.testing.assert_array_equal(cs.target_pc(2), np.full((20, 3), 2.0))
    assert cs.target_imgs().shape == (4, 8, 8, 4)
    assert datasets.natural_key("state_10.xz") > datasets.natural_key("state_9.xz")
