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

The gel-only drivers sim_gelonly.cpp (a 2x4x24 bilayer, anchored, pulse at
iteration 100) and sim_gelonly_tube.cpp (a 200x30x30 block hollowed into a
square tube, pulse at iteration 500) run for 1000 iterations too.  They read
device members back with raw runtime copies (m_drn and m_dvm with node-grid
strides, m_dVeln, m_dum) and print what they find every 500 / 1000
iterations; the same sequence through the C ABI must print the same lines
and write the same .dat files, and for sim_gelonly the printed tip
deflection and vm maximum must match the FP64 oracle.

sim_peristaltic.cpp (two 40x40x6 slabs in an 81x81x61 channel with
GEL_SIZE_X=40, pulse at iteration 500) runs the coupled step with two gels
for 1000 iterations: its status lines and files are compared the same way.

sim.cpp (two 20x20x4 gels in a 150x50x10 channel), sim_array.cpp (a 3x3
array of 20x20x4 gels, pulse delays drawn with srand(42)/rand()) and
sim_cilia.cpp (a 4x4 array of anchored 4x4x24 gels, the same delays) fix
their length at compile time; each runs until its status line of iteration
3000 and is stopped, and its status lines and the files of snapshots 0 and
1 must match the same loop through the C ABI.

The sim_single scene is then driven through the C ABI from Python with the
driver's call sequence -- per-object methods before iteration 500,
glsm_system_step (sim.cpp's loop body, the same calls) from there on -- and
the objects' writeFiles into a second directory.  Every .dat file must be
byte-identical.  The driver is built only where the reference tree is
mounted; the test fails if it is missing.
"""
import ctypes
import os
import queue
import re
import subprocess
import sys
import threading
import time

import numpy as np
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


# ---------------------------------------------------------------- gel only
def run_driver(name, run_dir, env_extra):
    exe = os.path.join(ROOT, "oracle", "_ref", name)
    assert os.path.exists(exe), f"{exe} is missing: build it with `make -C oracle drivers`"
    env = dict(os.environ)
    env.update(env_extra)
    print(f"running {name} {env_extra}", flush=True)
    r = subprocess.run([exe], cwd=str(run_dir), env=env, capture_output=True, text=True, timeout=300)
    sys.stdout.write(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr[-3000:]
    return r.stdout


def same_files(a_dir, b_dir, expected, skip=()):
    """Both directories hold exactly `expected`; every file but those in
    `skip` is non-empty and byte-identical (compared without handing the
    contents to the assertion, which would render them)."""
    a, b = sorted(os.listdir(a_dir)), sorted(os.listdir(b_dir))
    assert a == b and set(a) == expected, (a, b)
    for name in a:
        x, y = (a_dir / name).read_bytes(), (b_dir / name).read_bytes()
        if name in skip:
            print(f"{name}: not compared ({'identical' if x == y else 'differs'} in this run)", flush=True)
            continue
        if not x or x != y:
            pytest.fail(f"{name}: {len(x)} bytes from the driver, {len(y)} through the C ABI, "
                        f"first difference at byte {next((i for i, (p, q) in enumerate(zip(x, y)) if p != q), -1)}")


def gelonly_sequence(mod, runstep, out_dir=None):
    """sim_gelonly.cpp:41-125 through `mod` (the library or the oracle):
    the printed status lines, and the .dat files when out_dir is given."""
    size = (2, 4, 24)
    g = mod.Gel(size, (50.0, 50.0, size[2] / 2.0 + 3.0), gel_type=1, gel_id=1)
    g.set_anchor_z(3.0)
    g.reset_to_quiescent()
    LX, LY, LZ = (n + 1 for n in size)
    nN = (LX + 2) * (LY + 2) * (LZ + 2)
    lines, values = [], []
    for it in range(runstep + 1):
        if it == 100:
            g.fire_excitation_pulse()
        if it >= 100:
            g.step_chemistry(it)
            g.step_elasticity(it)
        g.record_center(it)
        if it % 500 == 0:
            rn = g.read("rn")
            # the driver copies (L+2)^3 doubles of m_dvm (sized to the node
            # array; the elements fill the front) and indexes them with
            # node-grid strides
            vm = np.zeros(nN)
            e = g.read("vm")
            vm[:e.size] = e
            base_x = tip_x = 0.0
            for yi in range(1, LY + 1):
                for xi in range(1, LX + 1):
                    base_x += rn[xi + yi * (LX + 2) + 1 * (LX + 2) * (LY + 2), 0]
                    tip_x += rn[xi + yi * (LX + 2) + LZ * (LX + 2) * (LY + 2), 0]
            base_x /= LX * LY
            tip_x /= LX * LY
            dx = tip_x - base_x
            vm_max = 0.0
            for zi in range(1, LZ + 1):
                for yi in range(1, LY + 1):
                    for xi in range(1, LX + 1):
                        v = vm[xi + yi * (LX + 2) + zi * (LX + 2) * (LY + 2)]
                        if not np.isnan(v) and v > vm_max:
                            vm_max = v
            lines.append("iter=%6d  tip_dx=%+.6f  vm_max=%.5f" % (it, dx, vm_max))
            values.append((dx, vm_max))
        if out_dir is not None:
            g.write_files(it)
    g.close()
    return lines, values


def test_sim_gelonly_driver_prints_and_writes_the_same(glsm, oracle, tmp_path):
    runstep = 1000
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    out = run_driver("sim_gelonly", run_dir, {"RUNSTEP": str(runstep)})
    printed = [ln for ln in out.splitlines() if ln.startswith("iter=")]
    mir = tmp_path / "mirror"
    mir.mkdir()
    cwd = os.getcwd()
    os.chdir(mir)
    try:
        lines, values = gelonly_sequence(glsm, runstep, out_dir=mir)
    finally:
        os.chdir(cwd)
    print("\n".join(lines), flush=True)
    assert printed == lines
    # the driver copies m_drn and m_dvm into the gel's host members at
    # iterations 500 and 1000 (sim_gelonly.cpp:76-101) while the writer
    # thread of iteration 0 may in principle still format them (gel.cu:543-553)
    same_files(run_dir / "data", mir,
               {f"gel1{k}{t}.dat" for k in ("rn", "Fn", "Veln", "rm", "um", "vm", "wm") for t in (0, 1)},
               skip={"gel1rn0.dat", "gel1vm0.dat"})
    # the FP64 oracle: the same deflection and chemistry maximum
    _, ref = gelonly_sequence(oracle, runstep)
    for (dx, vmax), (rdx, rvmax) in zip(values, ref):
        assert abs(dx - rdx) <= 1e-9 and abs(vmax - rvmax) <= 1e-9, (dx, rdx, vmax, rvmax)


def tube_sequence(G, runstep, wave_delay=500, wall=4):
    """sim_gelonly_tube.cpp:50-110 through the C ABI: status lines and files."""
    size = (200, 30, 30)
    g = G.Gel(size, (size[0] / 2.0, 15.0, 15.0), gel_type=1, gel_id=1, time=0)
    g.reset_to_quiescent()
    g.build_tube_mask(wall, False)
    LX, LY, LZ = g.node_grid
    lines = []
    for it in range(runstep + 1):
        if it == wave_delay:
            g.fire_excitation_pulse()
            lines.append(f"iter={it}: BZ excitation pulse fired")
        if it < wave_delay:
            g.record_center(it)
        else:
            g.step_elasticity(it)
            g.step_chemistry(it)
            g.record_center(it)
        if it % 1000 == 0:
            v = g.read("Veln")
            mag = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]).astype(np.float32)
            max_v = float(max(np.float32(0.0), mag.max()))
            um, vm = g.read("um"), g.read("vm")
            test_gi = g.get_index(1, 1, LZ // 2, 1)
            void_gi = g.get_index(1, LY // 2, LZ // 2, 1)
            lines.append("iter=%6d  Gel V_max = %.5f  um[solid]=%.5f vm[solid]=%.5f | um[void]=%.5f vm[void]=%.5f"
                         % (it, max_v, um[test_gi], vm[test_gi], um[void_gi], vm[void_gi]))
        g.write_files(it)
    g.close()
    return lines


def test_sim_gelonly_tube_driver_prints_and_writes_the_same(glsm, tmp_path):
    runstep = 1000
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    out = run_driver("sim_gelonly_tube", run_dir, {"RUNSTEP": str(runstep)})
    printed = [ln for ln in out.splitlines() if ln.startswith("iter=")]
    mir = tmp_path / "mirror"
    mir.mkdir()
    cwd = os.getcwd()
    os.chdir(mir)
    try:
        lines = tube_sequence(glsm, runstep)
    finally:
        os.chdir(cwd)
    print("\n".join(lines), flush=True)
    assert printed == lines
    # The driver copies m_dVeln, m_dum and m_dvm into the gel's host
    # members at iteration 1000 (sim_gelonly_tube.cpp:87-96) while the
    # writer thread of iteration 0 may still be formatting those members
    # (gel.cu:543-553 joins it only in the next writeFiles): the driver's
    # own race, present in the reference too, so those three files of time
    # 0 are listed but not compared.
    same_files(run_dir / "data_gelonly_tube", mir,
               {f"gel1{k}{t}.dat" for k in ("rn", "Fn", "Veln", "rm", "um", "vm", "wm") for t in (0, 1)},
               skip={"gel1Veln0.dat", "gel1um0.dat", "gel1vm0.dat"})


# ------------------------------------------------------------ two slabs
def peristaltic_sequence(G, runstep, gel_lx=40, wave_delay=500):
    """sim_peristaltic.cpp:56-140 through the C ABI: status lines and files."""
    size = (gel_lx, 40, 6)
    top = G.Gel(size, (gel_lx / 2.0, 20.0, 26.0), gel_type=1, gel_id=1)
    bot = G.Gel(size, (gel_lx / 2.0, 20.0, 4.0), gel_type=1, gel_id=2)
    top.reset_to_quiescent()
    bot.reset_to_quiescent()
    gels = [top, bot]
    c = G.Coupler(gels)
    f = G.Fluid((gel_lx, 40, 30), c)
    h = gel_lx / (f.grid[0] - 1)
    lines = []
    for it in range(runstep + 1):
        if it == wave_delay:
            top.fire_excitation_pulse()
            bot.fire_excitation_pulse()
            lines.append(f"iter={it}: BZ excitation pulse fired on both slabs")
        if it < wave_delay:
            for g in gels:
                g.record_center(it)
            c.pack_from_gels()
            f.step_velocity(it)
            f.step_concentration(it)
            c.apply_gel_repulsion()
            c.scatter_to_gels()
        else:
            G.system_step(gels, c, f, it)  # the same calls, in the driver's order
        if it % 1000 == 0:
            vg = np.float32(0.0)
            for g in gels:
                v = g.read("Veln")
                vg = max(vg, np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]).astype(np.float32).max())
            lines.append("iter=%6d  Vg_max=%.5f  lbm2gel=%.2f" % (it, float(vg), np.float32(f.nsub) / np.float32(h)))
        for g in gels:
            g.write_files(it)
        f.write_files(it)
    for o in gels + [f, c]:
        o.close()
    return lines


def test_sim_peristaltic_driver_prints_and_writes_the_same(glsm, tmp_path):
    runstep = 1000
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    out = run_driver("sim_peristaltic", run_dir, {"RUNSTEP": str(runstep), "GEL_SIZE_X": "40"})
    printed = [ln for ln in out.splitlines() if ln.startswith("iter=")]
    mir = tmp_path / "mirror"
    mir.mkdir()
    cwd = os.getcwd()
    os.chdir(mir)
    try:
        lines = peristaltic_sequence(glsm, runstep)
    finally:
        os.chdir(cwd)
    print("\n".join(lines), flush=True)
    assert printed == lines
    # as in the tube driver, the host copies of m_dVeln at iteration 1000
    # may overlap the writer threads of iteration 0 (gel.cu:543-553)
    gel_files = {f"gel{k}{n}{t}.dat" for k in (1, 2) for n in ("rn", "Fn", "Veln", "rm", "um", "vm", "wm")
                 for t in (0, 1)}
    same_files(run_dir / "data", mir, gel_files | {"Velb0.dat", "Conc0.dat"},
               skip={"gel1Veln0.dat", "gel2Veln0.dat"})


# ------------------------------------------- drivers of fixed length
# sim.cpp, sim_array.cpp and sim_cilia.cpp fix their length at compile time
# (`#define runstep` 150000 / 200000, sim.cpp:15, sim_array.cpp:17,
# sim_cilia.cpp:25).  Each runs until it prints its status line of
# iteration 3000 and is then stopped.  By then writeFiles(2000) has joined
# the writer threads of snapshot 1 (gel.cu:543-549, fluid.cu:144-146), so
# every file of snapshots 0 and 1 is complete; later ones may be partial
# and are not compared.
STOP_AT = 3000
SNAPSHOT = re.compile(r"^(?:gel\d+(?:rn|Fn|Veln|rm|um|vm|wm)|Velb|Conc)(\d+)\.dat$")


def run_driver_until(name, run_dir, last_iter, env_extra=None, timeout=240):
    """Start oracle/_ref/<name> in run_dir, collect its stdout until the
    status line of `last_iter`, then stop it; returns the lines read."""
    exe = os.path.join(ROOT, "oracle", "_ref", name)
    assert os.path.exists(exe), f"{exe} is missing: build it with `make -C oracle drivers`"
    env = dict(os.environ)
    env.update(env_extra or {})
    print(f"running {name} until iteration {last_iter}", flush=True)
    err = open(run_dir / "stderr.txt", "w")
    p = subprocess.Popen([exe], cwd=str(run_dir), env=env, stdout=subprocess.PIPE, stderr=err, text=True)
    lines, q = [], queue.Queue()

    def reader():
        for ln in p.stdout:
            q.put(ln.rstrip("\n"))
        q.put(None)

    t = threading.Thread(target=reader, daemon=True)
    t.start()
    deadline = time.monotonic() + timeout
    reached = False
    try:
        while True:
            ln = q.get(timeout=max(1.0, deadline - time.monotonic()))
            if ln is None:
                break
            lines.append(ln)
            m = re.match(r"iter=\s*(\d+)\s", ln)
            if m and int(m.group(1)) >= last_iter:
                reached = True
                break
    finally:
        if p.poll() is None:
            p.kill()
        p.wait(timeout=60)
        err.close()
    sys.stdout.write("\n".join(lines[-8:]) + "\n")
    assert reached, f"{name} stopped before iteration {last_iter}: " + (run_dir / "stderr.txt").read_text()[-3000:]
    return lines


def delayed_array_sequence(G, fluid_size, gel_size, centres, delays, last_iter, anchor_z=None):
    """The loop shared by sim.cpp:44-111, sim_array.cpp:66-133 and
    sim_cilia.cpp:73-135 through the C ABI: gel k (id k + 1) stays quiescent
    until its delay, when its excitation pulse fires; status lines every
    1000 iterations; files every iteration as the drivers write them."""
    gels = []
    for k, (pos, d) in enumerate(zip(centres, delays)):
        g = G.Gel(gel_size, pos, gel_type=1, gel_id=k + 1, time=0)
        if anchor_z is not None:
            g.set_anchor_z(anchor_z)
        if d > 0:
            g.reset_to_quiescent()
        gels.append(g)
    c = G.Coupler(gels)
    f = G.Fluid(fluid_size, c)
    h = fluid_size[0] / (f.grid[0] - 1)
    lines = []
    for it in range(last_iter + 1):
        for g, d in zip(gels, delays):
            if d > 0 and it == d:
                g.fire_excitation_pulse()
        for g, d in zip(gels, delays):
            if d > 0 and it < d:
                g.record_center(it)
                continue
            g.step_elasticity(it)
            g.step_chemistry(it)
            g.record_center(it)
        c.pack_from_gels()
        f.step_velocity(it)
        f.step_concentration(it)
        c.apply_gel_repulsion()
        c.scatter_to_gels()
        if it % 1000 == 0:
            vg = np.float32(0.0)
            for g in gels:
                v = g.read("Veln")
                vg = max(vg, np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]).astype(np.float32).max())
            lines.append("iter=%6d  Vg_max=%.5f  lbm2gel=%.2f" % (it, float(vg), np.float32(f.nsub) / np.float32(h)))
        for g in gels:
            g.write_files(it)
        f.write_files(it)
    for o in gels + [f, c]:
        o.close()
    return lines


def libc_delays(n, seed=42, period=50000.0):
    """(int)((double)rand() / RAND_MAX * T) for n gels after srand(seed), as
    sim_array.cpp:43-51 and sim_cilia.cpp:48-53 draw them (glibc rand)."""
    libc = ctypes.CDLL(None)
    libc.srand(ctypes.c_uint(seed))
    return [int(libc.rand() / 2147483647 * period) for _ in range(n)]


def compare_first_snapshots(drv_dir, mir_dir, ngel):
    """Files of snapshots 0 and 1 in both directories: the same names, and
    the same bytes except gel*Veln{0,1}.dat (the drivers copy m_dVeln into
    m_hVeln at every status line while that snapshot's writer thread may
    still format it, gel.cu:543-553: the drivers' own race)."""
    def first(d):
        return {n for n in os.listdir(d) if SNAPSHOT.match(n) and int(SNAPSHOT.match(n).group(1)) <= 1}
    expected = {f"gel{k}{n}{t}.dat" for k in range(1, ngel + 1) for n in ("rn", "Fn", "Veln", "rm", "um", "vm", "wm")
                for t in (0, 1)} | {"Velb0.dat", "Conc0.dat"}
    a, b = first(drv_dir), first(mir_dir)
    assert a == b == expected, (sorted(a ^ expected), sorted(b ^ expected))
    for name in sorted(a):
        x, y = (drv_dir / name).read_bytes(), (mir_dir / name).read_bytes()
        if "Veln" in name:
            continue
        if not x or x != y:
            pytest.fail(f"{name}: {len(x)} bytes from the driver, {len(y)} through the C ABI, "
                        f"first difference at byte {next((i for i, (p, q) in enumerate(zip(x, y)) if p != q), -1)}")


@pytest.mark.parametrize("name", ["sim", "sim_array", "sim_cilia"])
def test_fixed_length_driver_prints_and_writes_the_same(glsm, tmp_path, name):
    if name == "sim":  # sim.cpp:26-34, GEL2_DELAY unset
        scene = dict(fluid_size=(150, 50, 10), gel_size=(20, 20, 4), centres=[(50.0, 25.0, 5.0), (100.0, 25.0, 5.0)],
                     delays=[0, 0])
        env = {"GEL2_DELAY": "0"}
    elif name == "sim_array":  # sim_array.cpp:33-78
        scene = dict(fluid_size=(140, 140, 10), gel_size=(20, 20, 4),
                     centres=[(float(x), float(y), 7.0) for y in (30, 70, 110) for x in (30, 70, 110)],
                     delays=libc_delays(9))
        env = {"RAND_SEED": "42"}
    else:  # sim_cilia.cpp:36-84
        scene = dict(fluid_size=(100, 100, 30), gel_size=(4, 4, 24),
                     centres=[(float(x), float(y), 15.0) for y in (10, 30, 50, 70) for x in (10, 30, 50, 70)],
                     delays=libc_delays(16), anchor_z=3.0)
        env = {"RAND_SEED": "42"}
    print(f"{name}: delays {scene['delays']}", flush=True)
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    printed = [ln for ln in run_driver_until(name, run_dir, STOP_AT, env) if ln.startswith("iter=")]
    mir = tmp_path / "mirror"
    mir.mkdir()
    cwd = os.getcwd()
    os.chdir(mir)
    try:
        lines = delayed_array_sequence(glsm, last_iter=STOP_AT, **scene)
    finally:
        os.chdir(cwd)
    print("\n".join(lines), flush=True)
    assert printed == lines
    compare_first_snapshots(run_dir / "data", mir, len(scene["centres"]))
