"""The C ABI library loads and exports every entry point include/rwpt.h
declares (no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import pytest

from rwpt import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "rwpt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rwpt_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "rwpt_create" in names and "rwpt_advance" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(engine.library_path())
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol():
    assert set(_declared()) == set(engine.EXPORTS)


def test_default_options_are_reference_defaults():
    o = engine.default_options()
    assert o.advection_kind == 2 and o.time_step_kind == 4
    assert list(o.dimension_mask) == [1, 1, 1]
    assert o.base_seed[0] == 0x1DADBEEFBAADD0D0
    assert o.base_seed[1] == 0x5BADD0D0DEADBEEF


def test_create_rejects_bad_arguments_without_touching_the_device():
    lib = engine.lib()
    h = C.c_void_p()
    o = engine.default_options(dimension_mask=(0, 0, 0))
    g = engine.RwptGrid()
    # null grid pointers and a zero dimension mask are rejected on the host
    assert lib.rwpt_create(C.byref(g), C.byref(o), C.byref(h)) == -1
    assert lib.rwpt_create(None, C.byref(o), C.byref(h)) == -1


def test_tuning_rejects_block_above_launch_bound():
    # rwpt_set_tuning validates on the host; a null engine is rejected first
    lib = engine.lib()
    assert lib.rwpt_set_tuning(None, 64, 512, 1, 1) == -1
