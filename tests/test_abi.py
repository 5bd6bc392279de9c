"""The C ABI: the library loads, exports every function include/myrt.h declares, and reports
errors through status codes and mrt_last_error() (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import myrt_amd as m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(mrt_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported(ensure_built):
    lib = ctypes.CDLL(m.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # The Python binding covers exactly the declared functions.
    assert set(m.SIGNATURES) == names


def test_invalid_arguments_fail_cleanly(ensure_built):
    s = m.Scene()
    with pytest.raises(m.MrtError, match="out of range"):
        s.push_geometry(np.array([0, 1, 5], np.uint32), np.zeros((3, 3), np.float32))
    with pytest.raises(m.MrtError, match="multiple of 3"):
        s.push_geometry(np.array([0, 1], np.uint32), np.zeros((3, 3), np.float32))
    with pytest.raises(m.MrtError, match="unknown geometry"):
        s.enqueue(3)
    g = s.push_geometry(np.array([0, 1, 2], np.uint32), np.eye(3, dtype=np.float32))
    with pytest.raises(m.MrtError, match="not invertible"):
        s.enqueue(g, np.zeros((4, 4), np.float32))
    with pytest.raises(m.MrtError, match="before mrt_scene_prepare"):
        s.trace_closest(m.make_rays([[0, 0, 0]], [[0, 0, 1]]))
    with pytest.raises(m.MrtError, match="at least one box"):
        m.bvh_build(np.zeros(0, m.AABB_DTYPE))


def test_prepare_consumes_drawables_on_host(ensure_built):
    """scene::prepare takes this frame's drawables; a frame with none keeps the previous one."""
    s = m.Scene()
    g = s.push_geometry(np.array([0, 1, 2], np.uint32), np.eye(3, dtype=np.float32))
    s.enqueue(g)
    s.enqueue(g, np.diag([2, 2, 2, 1]).astype(np.float32))
    r = s.build_host()
    assert r["num_instances"] == 2 and r["drawables_changed"] == 1 and r["tlas_wide_nodes"] >= 1
    r = s.build_host()
    assert r["num_instances"] == 2
    s.enqueue(g)
    r = s.build_host()
    assert r["num_instances"] == 1
