"""The reference's own driver runs on the library.

oracle/_ref/sim_single is the reference's src/sim_single.cpp built against
this library's C++ boundary (oracle/Makefile `drivers`: the CUDA-runtime ->
HIP-runtime rename, nothing else).  Its scene is configured from the
environment (sim_single.cpp:22-28, 40-55): RUNSTEP, FLUID_SIZE_X/Y/Z,
GEL_SIZE_X/Y/Z.  The test runs it for 1001 iterations, so the loop
(sim_single.cpp:83-120) covers the chemistry-free phase before the
excitation pulse at iteration 500, the pulse, the coupled phase, the gel
output at iterations 0 and 1000 (gel1{rn,Fn,Veln,rm,um,vm,wm}<t>.dat, t = 0
and 1) and the fluid output at iteration 0 (Velb0.dat, Conc0.dat).  (The
centre-ring file gel1bodycenter.dat is written only when int(iter * dt) is
a multiple of 1000, i.e. from iteration 10^6 on, gel.cu:505; its layout is
tested in test_gpu_io.py.)

The same scene is then driven through the C ABI from Python with the
driver's call sequence -- per-object methods before iteration 500,
glsm_system_step (sim.cpp's loop body, the same calls) from there on -- and
the objects' writeFiles into a second directory.  Every .dat file must be
byte-identical.  The driver is built only where the reference tree is
mounted; the test fails if it is missing.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "sim_single")
RUNSTEP, FIRE = 1000, 500
FLUID, GEL = (20, 20, 30), (2, 4, 24)


def scene_env():
    env = dict(os.environ)
    for k in ("IBM_BETA", "IBM_RAMP_STEPS", "GLSM_FLUID_STORAGE"):
        env.pop(k, None)
    env.update(RUNSTEP=str(RUNSTEP), FLUID_SIZE_X=str(FLUID[0]), FLUID_SIZE_Y=str(FLUID[1]),
               FLUID_SIZE_Z=str(FLUID[2]), GEL_SIZE_X=str(GEL[0]), GEL_SIZE_Y=str(GEL[1]), GEL_SIZE_Z=str(GEL[2]))
    return env


def mirror(G, out_dir):
    """sim_single.cpp:63-120 through the C ABI."""
    cwd = os.getcwd()
    os.makedirs(out_dir, exist_ok=True)
    os.chdir(out_dir)
    try:
        g = G.Gel(GEL, (FLUID[0] / 2.0, FLUID[1] / 2.0, 15.0), gel_type=1, gel_id=1)
        g.set_anchor_z(3.0)
        g.reset_to_quiescent()
        c = G.Coupler([g])
        f = G.Fluid(FLUID, c)
        for it in range(RUNSTEP + 1):
            if it == FIRE:
                g.fire_excitation_pulse()
            if it < FIRE:
                g.record_center(it)
                c.pack_from_gels()
                f.step_velocity(it)
                f.step_concentration(it)
                c.apply_gel_repulsion()
                c.scatter_to_gels()
            else:
                G.system_step([g], c, f, it)
            g.write_files(it)
            f.write_files(it)
        for o in (g, f, c):
            o.close()  # joins the writer threads
    finally:
        os.chdir(cwd)


def test_sim_single_driver_writes_the_same_files(glsm, tmp_path):
    assert os.path.exists(DRIVER), f"{DRIVER} is missing: build it with `make -C oracle drivers`"
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    print(f"running {DRIVER} RUNSTEP={RUNSTEP} fluid {FLUID} gel {GEL}", flush=True)
    r = subprocess.run([DRIVER], cwd=str(run_dir), env=scene_env(), capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert f"iter={RUNSTEP:6d}" in r.stdout
    mirror(glsm, str(tmp_path / "mirror"))
    drv = sorted(os.listdir(run_dir / "data"))
    mir = sorted(os.listdir(tmp_path / "mirror"))
    print("files:", drv, flush=True)
    assert drv == mir
    expected = {f"gel1{k}{t}.dat" for k in ("rn", "Fn", "Veln", "rm", "um", "vm", "wm") for t in (0, 1)}
    assert expected | {"Velb0.dat", "Conc0.dat"} == set(drv)
    for name in drv:
        a = (run_dir / "data" / name).read_bytes()
        b = (tmp_path / "mirror" / name).read_bytes()
        assert len(a) > 0 and a == b, name
