"""The reference's drivers compile against this library's C++ boundary.

Each src/sim*.cpp of the reference is read as text (it is not copied into
this repository), the mechanical CUDA-runtime -> HIP-runtime rename is
applied -- the only edit the boundary needs, since no CUDA compatibility
header is provided -- and the result is compiled and linked with hipcc
against include/{gel,coupling,fluid}.h and libglsm_fluid.so.  The test also
checks that the rename touched nothing but runtime identifiers, i.e. every
line that uses Gel / Coupler / Fluid compiles unchanged.  Skipped where the
reference tree is not mounted.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
DRIVERS = ["sim.cpp", "sim_single.cpp", "sim_array.cpp", "sim_cilia.cpp", "sim_peristaltic.cpp",
           "sim_gelonly.cpp", "sim_gelonly_tube.cpp"]
RENAMES = [("<cuda_runtime.h>", "<hip/hip_runtime.h>"), ("cudaSetDevice", "hipSetDevice"),
           ("cudaDeviceSynchronize", "hipDeviceSynchronize"), ("cudaMemcpyDeviceToHost", "hipMemcpyDeviceToHost"),
           ("cudaMemcpy", "hipMemcpy"), ("cudaGetLastError", "hipGetLastError"),
           ("cudaGetErrorString", "hipGetErrorString"), ("cudaError_t", "hipError_t"), ("cudaSuccess", "hipSuccess")]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_SRC) or shutil.which("hipcc") is None,
                                reason="reference sources or hipcc not available")


def renamed(text):
    for a, b in RENAMES:
        text = text.replace(a, b)
    return text


@pytest.fixture(scope="module")
def library():
    import glsm_fluid_3d_amd as G

    if not os.path.exists(G.LIB_PATH):
        G.compile_library()
    return G.LIB_PATH


@pytest.mark.parametrize("driver", DRIVERS)
def test_reference_driver_builds_against_boundary(library, driver, tmp_path):
    src = open(os.path.join(REF_SRC, driver), encoding="utf-8", errors="replace").read()
    out = renamed(src)
    # only runtime identifiers changed
    def strip_runtime(line, pre):
        line = re.sub(r"<cuda_runtime\.h>|<hip/hip_runtime\.h>", "<RT>", line)
        return re.sub(r"\b" + pre + r"[A-Z]\w*", "RT", line)

    changed = [(a, b) for a, b in zip(src.splitlines(), out.splitlines()) if a != b]
    assert len(src.splitlines()) == len(out.splitlines())
    for a, b in changed:
        assert strip_runtime(a, "cuda") == strip_runtime(b, "hip"), (a, b)
    assert "cuda" not in out
    path = tmp_path / driver
    path.write_text(out)
    exe = tmp_path / driver.replace(".cpp", "")
    libdir = os.path.dirname(library)
    r = subprocess.run(["hipcc", "-O1", "-std=c++17", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
                        str(path), "-L", libdir, "-lglsm_fluid", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert exe.exists()
