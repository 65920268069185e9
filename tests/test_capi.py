"""The C ABI library (CPU-side checks, no kernels launched): it loads, it
exports every entry point include/glsm_fluid.h declares, and its error path
reports instead of crashing when no GPU is usable."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "glsm_fluid.h")


def declared_entry_points():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(glsm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    import glsm_fluid_3d_amd as G

    path = G.LIB_PATH
    if not os.path.exists(path):
        G.compile_library()
    return C.CDLL(path), G


def test_header_declares_the_class_api():
    names = declared_entry_points()
    for must in ["glsm_gel_create", "glsm_gel_step_elasticity", "glsm_gel_step_chemistry", "glsm_gel_record_center",
                 "glsm_gel_write_files", "glsm_gel_reset_to_quiescent", "glsm_gel_fire_excitation_pulse",
                 "glsm_gel_set_anchor_z", "glsm_gel_build_tube_mask", "glsm_coupler_create",
                 "glsm_coupler_pack_from_gels", "glsm_coupler_scatter_to_gels", "glsm_coupler_apply_gel_repulsion",
                 "glsm_fluid_create", "glsm_fluid_step_velocity", "glsm_fluid_step_concentration",
                 "glsm_fluid_write_files", "glsm_system_step"]:
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    L, G = lib
    missing = [n for n in declared_entry_points() if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_entry_points()) == set(G.SIGNATURES)


def test_version_and_error_reporting_without_device(lib):
    L, G = lib
    L.glsm_version.restype = C.c_char_p
    assert b"gfx950" in L.glsm_version()
    L.glsm_last_error.restype = C.c_char_p
    # a null handle is rejected with EINVAL and a message, before any HIP call
    rc = L.glsm_gel_step_chemistry(None, 0)
    assert rc == -1
    assert b"null" in L.glsm_last_error()
    rc = L.glsm_fluid_step_velocity(None, 0)
    assert rc == -1
