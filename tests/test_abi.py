"""CPU checks of the C ABI and host logic (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from mahakala_amd import _lib
from mahakala_amd._array import MkArray, wrap


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (mk_\w+)", out))
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    for s in declared:
        assert hasattr(L, s)
    assert L.mk_version() >= 10000


def test_argument_errors_are_reported_without_a_device():
    L = _lib.lib()
    rc = L.mk_camera_grid(0.5, 30.0, 1000.0, -10.0, 10.0, 0, None, None)
    assert rc == -1
    assert b"mk_camera_grid" in L.mk_last_error()
    p = _lib.mk_render_params()
    assert L.mk_render(None, C.byref(p), None, None, None) == -1
    with pytest.raises(_lib.MkError):
        _lib.check(L.mk_integrate_history(None, 1, 10, 40.0, 1e-4, 0.5, 5, None, None, None))


def test_render_params_layout_matches_header():
    text = open(_lib.HEADER_PATH).read()
    block = text[text.index("typedef struct mk_render_params"):text.index("} mk_render_params;")]
    fields = re.findall(r"^\s*(double|int32_t)\s+(\w+);", block, re.M)
    assert [n for _, n in fields] == [n for n, _ in _lib.mk_render_params._fields_]
    assert C.sizeof(_lib.mk_render_params) == 16 * 8 + 4 * 4


def test_mkarray_at_is_functional():
    a = wrap(np.arange(6.0))
    b = a.at[a > 2].set(0.0)
    assert isinstance(b, MkArray)
    assert np.array_equal(a, np.arange(6.0))
    assert np.array_equal(b, [0, 1, 2, 0, 0, 0])
    c = b.at[np.array([0, 1])].add(10.0)
    assert np.array_equal(c, [10, 11, 2, 0, 0, 0])
    assert np.array_equal(a.at[1:3].get(), [1, 2])
    assert np.array_equal(a.at[a > 4].multiply(2.0), [0, 1, 2, 3, 4, 10])


def test_mahakala_alias_exposes_reference_surface():
    import mahakala as ma
    import mahakala.geodesics as mg
    import mahakala_amd.geodesics as ag
    from mahakala.grmhd.athenak import AthenakFluidModel  # noqa: F401
    from mahakala.images import make_image  # noqa: F401
    from mahakala.electrons import rlow_rhigh_model  # noqa: F401
    from mahakala.constants import Msun, EE, KB, CL, ME, HPL, GNEWT  # noqa: F401
    assert mg is ag
    assert set(ma.__all__) == {"find_shadow_bisection", "find_shadow_bisection_angles", "geodesic_integrator",
                               "initialize_geodesics_at_camera", "synchrotron_coefficients",
                               "solve_specific_intensity", "solve_attenuated_emissivity"}
    for name in ("metric", "imetric", "radius_cal", "radius_EH", "RK4_gen", "rhs", "initial_condition",
                 "get_initial_grid", "get_camera_pixel", "select_photons_integrator"):
        assert callable(getattr(mg, name))


def test_server_fails_loudly_without_a_device():
    L = _lib.lib()
    if L.mk_device_info(None, 0, None, None) == 0:
        pytest.skip("a GPU is visible")
    from mahakala_amd.serving import Server
    with pytest.raises(_lib.MkError):
        Server("mixed")
    with pytest.raises(ValueError):
        Server("fp16")
    cfg = _lib.mk_server_config(3, 0, 0, 0, 0, 0, 0)
    h = _lib.vp()
    assert L.mk_server_create(C.byref(cfg), C.byref(h)) == -1


def test_server_structs_match_header():
    text = open(_lib.HEADER_PATH).read()
    for name in ("mk_server_config", "mk_image_request", "mk_shadow_request", "mk_server_stats"):
        block = text[text.index(f"typedef struct {name}"):text.index(f"}} {name};")]
        block = re.sub(r"/\*.*?\*/", "", block, flags=re.S)
        fields = []
        for m in re.finditer(r"^\s*(?:const\s+)?(?:double|int32_t|int64_t|size_t)\*?\s+([\w,\s*]+);", block, re.M):
            fields += [f.strip().lstrip("*") for f in m.group(1).split(",")]
        assert fields == [n for n, _ in getattr(_lib, name)._fields_], name


def test_athdf_library_exports_every_declared_symbol():
    from mahakala_amd import _athdf
    if not os.path.exists(_athdf.LIB_PATH):
        pytest.skip("libmk_athdf.so not built (no HDF5 headers)")
    text = open(_athdf.HEADER_PATH).read()
    declared = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(mk_athdf_\w+)\s*\(", text, re.M)))
    assert len(declared) == 6
    out = subprocess.run(["nm", "-D", "--defined-only", _athdf.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (mk_athdf_\w+)", out))
    assert not [s for s in declared if s not in exported]
    _athdf.lib()


def test_header_is_plain_c():
    """include/mahakala_hip.h is a C header (the boundary a cgo, JNI or ctypes
    binding includes or restates): it compiles as strict C99 and as C++, and a
    C program that uses its structs and flags links against the library."""
    import shutil
    import tempfile
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = os.path.join(root, "include")
    src = ("#include <stddef.h>\n#include <string.h>\n#include \"mahakala_hip.h\"\n"
           "int main(void) {\n"
           "  mk_render_params p;\n  mk_server_config cfg;\n  mk_image_request r;\n  mk_shadow_request sh;\n"
           "  memset(&p, 0, sizeof p); memset(&cfg, 0, sizeof cfg); memset(&r, 0, sizeof r);\n"
           "  memset(&sh, 0, sizeof sh);\n"
           "  p.flags = MK_RENDER_CLASSIC | MK_RENDER_STATIC_TILES | MK_RENDER_ROW_MAJOR;\n"
           "  cfg.flags = MK_SERVER_PERSISTENT | MK_SERVER_SHADOW_IN_LAUNCH | MK_SERVER_IMAGE_IN_LAUNCH;\n"
           "  r.priority = 1; sh.priority = 2;\n"
           "  return mk_version() > 0 && p.flags && cfg.flags && r.priority + sh.priority == 3 ? 0 : 1;\n"
           "}\n")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        with open(c, "w") as f:
            f.write(src)
        for cmd in (["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-I", inc, "-c", c,
                     "-o", os.path.join(d, "t.o")],
                    ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", inc, "-x", "c++", "-c", c,
                     "-o", os.path.join(d, "tpp.o")]):
            r = subprocess.run(cmd, capture_output=True, text=True)
            assert r.returncode == 0, r.stderr
        lib_dir = os.path.dirname(_lib.LIB_PATH)
        if os.path.exists(_lib.LIB_PATH):
            r = subprocess.run(["gcc", "-o", os.path.join(d, "t"), os.path.join(d, "t.o"), "-L", lib_dir,
                                "-lmahakala_hip", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr
