"""CPU checks of the C ABI: libmhdp.so loads, exports every entry point include/mhdp.h
declares, and its host-only functions (anchors(), hierarchy()) agree with the oracle.
No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from mhd_simulator_amd import _lib, fluid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mhdp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhdp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SYMBOLS)


def test_library_is_a_gfx950_hip_build():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(tmp_path):
    saved = _lib._lib
    _lib._lib = None
    try:
        with pytest.raises(RuntimeError, match="missing"):
            _lib.load(str(tmp_path / "nope.so"))
    finally:
        _lib._lib = saved


@pytest.mark.parametrize("seed", range(8))
def test_library_anchors_match_oracle(orc, seed):
    rng = np.random.default_rng(100 + seed)
    dims = tuple(int(v) for v in rng.integers(1, 12, size=3))
    periodic = tuple(bool(v) for v in rng.integers(0, 2, size=3))
    f = (rng.random(int(np.prod(dims))) < rng.uniform(0.3, 0.9)).astype(np.uint8)
    a, ca = fluid.anchors(f, dims, periodic)
    b, cb = orc.anchors(f, dims, periodic)
    assert ca == cb and np.array_equal(a, b)


def test_library_hierarchy_known_answers():
    for dims, expected in [((256, 256, 256, 256 ** 3), 2396736), ((33, 1, 1, 33), 34), ((1, 1, 1, 1), 0)]:
        levels, coarse = fluid.hierarchy(dims)
        assert coarse == expected and all(d <= 4 for d in levels[-1][:3])
    assert fluid.hierarchy((0, 1, 1, 0)) is None
    assert fluid.hierarchy((2 ** 32 - 1, 1, 1, 2 ** 32 - 1)) is None
    assert fluid.hierarchy((2 ** 32 - 1,) * 4) is None
    assert len(fluid.hierarchy((512, 512, 512, 512 ** 3))[0]) == 8


def test_config_validation_mirrors_reference():
    fluid.FluidConfig().validate()
    for bad in (dict(relative_tolerance=1e-8), dict(relative_tolerance=0.1), dict(max_iterations=0),
                dict(max_iterations=20001), dict(absolute_divergence_tolerance=-1.0)):
        with pytest.raises(ValueError, match="Invalid fluid timestep"):
            fluid.FluidConfig(**bad).validate()
