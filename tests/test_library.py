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


CPP_PROGRAM = r"""
#include "mikroptim_hip.hpp"
#include <cstdio>
int main() {
    int fails = 0;
    try {
        mko_hip::check(MKO_ERR_INVALID);
        ++fails;
    } catch (const std::invalid_argument&) {
    }
    try {
        mko_hip::check(MKO_ERR_DEVICE);
        ++fails;
    } catch (const mko_hip::DeviceError&) {
    }
    mko_variables v{};
    v.count = 3;
    const double x[4] = {0, 0, 0, 0};
    try {
        mko_hip::TraceEngine::runs_of(std::span<const double>(x, 4), v);
        ++fails;
    } catch (const std::invalid_argument&) {
    }
    if (mko_hip::TraceEngine::runs_of(std::span<const double>(x, 3), v) != 1) ++fails;
    // without a device the engine reports a device error; with one it is created
    try {
        mko_hip::TraceEngine e(0);
        if (!e.handle()) ++fails;
        std::puts("engine");
    } catch (const mko_hip::DeviceError& e) {
        std::puts("no device");
    }
    std::printf("fails=%d\n", fails);
    return fails;
}
"""


def test_cpp_header_builds_links_and_maps_errors(tmp_path):
    """include/mikroptim_hip.hpp compiles on its own with the host compiler, links against the
    library and maps the C-ABI's error codes to the reference's exception types."""
    import subprocess

    src = tmp_path / "use_header.cpp"
    src.write_text(CPP_PROGRAM)
    exe = tmp_path / "use_header"
    libdir = os.path.dirname(engine.LIB_PATH)
    subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", libdir, "-l:" + os.path.basename(engine.LIB_PATH),
                    "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "fails=0" in out.stdout


def test_reference_adapter_variables_view():
    """mko::hip::VariablesView (include/mikroptim_hip_reference.hpp) turns the reference's
    VariableSet into the library's variable description: the same types, batch surfaces and glass
    components as the description the VariableSet was built from, the same typical sizes and,
    where that description has finite bounds, the same bounds."""
    import numpy as np
    import pytest

    from mikroptim_amd import systems
    from oracle_lib import reference

    ref = reference()
    if ref is None:
        pytest.skip("reference build absent")
    # uneven topologies 1 and 5: whole glass triples (the reference's glass entries expand to
    # all three components)
    for b, v in (systems.objective(variants=2, seed=2), systems.singlet(2, seed=1),
                 systems.uneven_topology(1, variants=2), systems.uneven_topology(5, variants=2)):
        assert (v.type == 3).sum() % 3 == 0
        view = ref.variables_view(b, v)
        assert view["count"] == v.size
        assert np.array_equal(view["type"], v.type)
        assert np.array_equal(view["surface"], v.batch_surface)
        if v.has_glass:
            assert np.array_equal(view["component"], v.component)
        glass = v.type == 3
        fin = np.isfinite(v.lower) & ~glass
        assert np.array_equal(view["lower"][fin], v.lower[fin])
        fin = np.isfinite(v.upper) & ~glass
        assert np.array_equal(view["upper"][fin], v.upper[fin])
        assert np.array_equal(view["typical"][~glass], v.typical[~glass])
