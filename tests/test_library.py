"""The C-ABI library loads on a CPU-only host and exports every symbol the headers declare."""

import ctypes
import os
import re

from mikroptim_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mikroptim_hip.h")).read()
    return sorted(set(re.findall(r"\b(mko_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = engine.library()
    names = declared_symbols()
    assert set(names) == set(engine.EXPORTS)
    for name in names:
        assert hasattr(lib, name), name


def test_library_is_built_for_gfx950():
    data = open(engine.LIB_PATH, "rb").read()
    assert b"gfx950" in data
