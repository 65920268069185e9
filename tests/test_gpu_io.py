"""Output layout, text restart and binary snapshots of the HIP path.

  * .dat layout (gel.cu:452-541, fluid.cu:115-138): every file the writers
    produce is compared byte for byte with the layout rendered here from the
    fields read through the C ABI -- `setw(9) << to_string(x)` (printf "%f",
    right-aligned in 9 columns) joined by six spaces for the 3-column files,
    and `setw(9) << double` (printf "%g", 6 significant digits) for vm / wm;
    real nodes (1..L) and interior elements (1..L-1) in z, y, x loop order.
  * text restart (gel.cu:221-282, fluid.cu:57-83): a gel built with time=t
    holds exactly the values parsed from gel<id>{rn,um,vm,wm}<t>.dat; a fluid
    built with time=t holds u and c parsed from Velb<t>.dat / Conc<t>.dat and
    populations at rest (f_q = w_q), the reference's lossy restart.
  * binary snapshots: k steps, save, fresh objects, load, k more steps equal
    2k uninterrupted steps bit for bit, in every storage mode.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)


def small_system(G, storage="f32", seed=0, n=6):
    g = G.Gel((n, n, n), (10.0, 10.0, 10.0), gel_id=1)
    g.perturb_chemistry(seed, 0.01)
    c = G.Coupler([g])
    f = G.Fluid((20, 20, 20), c)
    f.set_periodic(1, 1, 1)
    f.set_ibm(1.0, 20.0)  # short ramp so the gel drives the fluid within a few steps
    f.set_storage(storage)
    return g, c, f


def close_all(*objs):
    for o in objs:
        o.close()


def ts(x):  # std::to_string(double)
    return "%f" % x


def row3(v):
    return "      ".join("%9s" % ts(a) for a in v) + "\n"


def real_nodes(g, arr):
    L = g.node_grid[0]
    a = np.asarray(arr).reshape((L + 2, L + 2, L + 2) + np.asarray(arr).shape[1:])
    return a[1:L + 1, 1:L + 1, 1:L + 1].reshape((-1,) + a.shape[3:])


def interior_elements(g, arr):
    L = g.node_grid[0]
    a = np.asarray(arr).reshape((L + 1, L + 1, L + 1) + np.asarray(arr).shape[1:])
    return a[1:L, 1:L, 1:L].reshape((-1,) + a.shape[3:])


def test_dat_layout_byte_exact(glsm, tmp_path, monkeypatch):
    G = glsm
    monkeypatch.chdir(tmp_path)
    g, c, f = small_system(G)
    for it in range(1, 8):
        G.system_step([g], c, f, it)
    fields = {k: g.read(k) for k in ("rn", "Fn", "Veln", "rm", "um", "vm", "wm")}
    u, conc = f.read("u"), f.read("c")
    g.write_files(1000)   # t = int(1000 * 1e-3) = 1
    f.write_files(2000)   # t = int(2000 * 1e-3f) = 2
    close_all(g, f, c)    # joins the writer threads
    expect = {
        "gel1rn1.dat": "".join(row3(v) for v in real_nodes(g, fields["rn"])),
        "gel1Fn1.dat": "".join(row3(v) for v in real_nodes(g, fields["Fn"])),
        "gel1Veln1.dat": "".join(row3(v) for v in real_nodes(g, fields["Veln"])),
        "gel1rm1.dat": "".join(row3(v) for v in interior_elements(g, fields["rm"])),
        "gel1um1.dat": "".join("%9s\n" % ts(v) for v in interior_elements(g, fields["um"])),
        "gel1vm1.dat": "".join("%9s\n" % ("%g" % v) for v in interior_elements(g, fields["vm"])),
        "gel1wm1.dat": "".join("%9s\n" % ("%g" % v) for v in interior_elements(g, fields["wm"])),
        "Velb2.dat": "".join(row3(v.astype(np.float64)) for v in u),
        "Conc2.dat": "".join("%9s\n" % ts(float(v)) for v in conc),
    }
    got = sorted(os.listdir(tmp_path))
    assert got == sorted(expect), got
    L = g.node_grid[0]
    n_lines = {"Velb2.dat": f.N, "Conc2.dat": f.N, "gel1rn1.dat": L ** 3, "gel1Fn1.dat": L ** 3,
               "gel1Veln1.dat": L ** 3}
    for name, text in expect.items():
        data = open(tmp_path / name).read()
        assert data.count("\n") == n_lines.get(name, (L - 1) ** 3), name
        assert data == text, name
    # the velocity snapshot is not all zeros (the gel drove the fluid)
    assert np.abs(u).max() > 0


def test_text_restart_reads_snapshot(glsm, tmp_path, monkeypatch):
    G = glsm
    monkeypatch.chdir(tmp_path)
    g, c, f = small_system(G)
    for it in range(1, 8):
        G.system_step([g], c, f, it)
    g.write_files(1000)
    f.write_files(2000)
    close_all(g, f, c)

    g2 = G.Gel((6, 6, 6), (10.0, 10.0, 10.0), gel_id=1, time=1)
    c2 = G.Coupler([g2])
    f2 = G.Fluid((20, 20, 20), c2, time=2)
    try:
        rn_file = np.loadtxt(tmp_path / "gel1rn1.dat")
        np.testing.assert_array_equal(real_nodes(g2, g2.read("rn")), rn_file)
        for k in ("um", "vm", "wm"):
            np.testing.assert_array_equal(interior_elements(g2, g2.read(k)), np.loadtxt(tmp_path / f"gel1{k}1.dat"))
        u_file = np.loadtxt(tmp_path / "Velb2.dat", dtype=np.float64).astype(np.float32)
        np.testing.assert_array_equal(f2.read("u"), u_file)
        np.testing.assert_array_equal(f2.read("c"), np.loadtxt(tmp_path / "Conc2.dat").astype(np.float32))
        pops = f2.read("f").reshape(19, -1)
        np.testing.assert_allclose(pops, np.broadcast_to(W[:, None], pops.shape), rtol=0, atol=1e-7)
    finally:
        close_all(g2, f2, c2)


@pytest.mark.parametrize("storage", ["f32", "f32full", "f16", "bf16"])
def test_binary_snapshot_resume_is_exact(glsm, tmp_path, storage):
    G = glsm
    g, c, f = small_system(G, storage)
    for it in range(0, 12):
        G.system_step([g], c, f, it)
    ref_gel = {k: g.read(k) for k in G.GEL_FIELDS}
    ref_u, ref_c, ref_f = f.read("u"), f.read("c"), f.read("fpost")
    close_all(g, f, c)

    g, c, f = small_system(G, storage)
    for it in range(0, 6):
        G.system_step([g], c, f, it)
    g.save_state(str(tmp_path / "gel.snap"))
    f.save_state(str(tmp_path / "fluid.snap"))
    close_all(g, f, c)

    g, c, f = small_system(G, "f32", seed=99)  # different state and storage before the load
    try:
        for it in range(0, 3):
            G.system_step([g], c, f, it)
        g.load_state(str(tmp_path / "gel.snap"))
        f.load_state(str(tmp_path / "fluid.snap"))
        for it in range(6, 12):
            G.system_step([g], c, f, it)
        for k in G.GEL_FIELDS:
            np.testing.assert_array_equal(g.read(k), ref_gel[k], err_msg=k)
        np.testing.assert_array_equal(f.read("u"), ref_u)
        np.testing.assert_array_equal(f.read("c"), ref_c)
        np.testing.assert_array_equal(f.read("fpost"), ref_f)
    finally:
        close_all(g, f, c)


def test_snapshot_rejects_other_shapes(glsm, tmp_path):
    G = glsm
    g, c, f = small_system(G)
    g.save_state(str(tmp_path / "gel.snap"))
    f.save_state(str(tmp_path / "fluid.snap"))
    g4 = G.Gel((4, 4, 4), (10.0, 10.0, 10.0), gel_id=1)
    c4 = G.Coupler([g4])
    f4 = G.Fluid((10, 10, 10), c4)
    try:
        with pytest.raises(G.GlsmError, match="differ"):
            g4.load_state(str(tmp_path / "gel.snap"))
        with pytest.raises(G.GlsmError, match="differ"):
            f4.load_state(str(tmp_path / "fluid.snap"))
        with pytest.raises(G.GlsmError, match="cannot open"):
            g.load_state(str(tmp_path / "absent.snap"))
        # a snapshot of another format version (version 1 stored one plane of
        # N values per population, not the block layout) is refused, not misread
        raw = bytearray(open(tmp_path / "fluid.snap", "rb").read())
        assert raw[:8] == b"GLSMSNAP" and int.from_bytes(raw[8:12], "little") == 2
        raw[8:12] = (1).to_bytes(4, "little")
        open(tmp_path / "fluid_v1.snap", "wb").write(bytes(raw))
        with pytest.raises(G.GlsmError, match="not a snapshot file of this version"):
            f.load_state(str(tmp_path / "fluid_v1.snap"))
    finally:
        close_all(g, f, c, g4, f4, c4)


def test_bodycenter_layout_byte_exact(glsm, tmp_path, monkeypatch):
    """gel<id>bodycenter.dat (gel.cu:505-536): written by the output of a
    time that is a multiple of 1000 (iteration 10^6 at dt = 1e-3) once an
    earlier output has set the gel's flag; 1000 rows, one per ring slot,
    `setw(9) << to_string(...)` of (time - 1000 + i, rm, Fn, Veln, vm, wm)
    joined by six spaces.  The ring slots are the centre records of
    iterations that are multiples of 1000 (gel.cu:706-713)."""
    G = glsm
    monkeypatch.chdir(tmp_path)
    g, c, f = small_system(G)
    for it in range(0, 6):
        G.system_step([g], c, f, it)
        g.record_center(1000 * (it + 1))  # ring slots 1..6 from the evolving state
    ring = {k: g.read(k) for k in ("center_rm", "center_Fn", "center_Veln", "center_vm", "center_wm")}
    g.write_files(0)        # first output: sets the flag
    g.write_files(1000000)  # t = 1000: writes the ring
    close_all(g, f, c)
    rows = []
    for i in range(1000):
        vals = [*ring["center_rm"][i], *ring["center_Fn"][i], *ring["center_Veln"][i], ring["center_vm"][i],
                ring["center_wm"][i]]
        rows.append("      ".join(["%9s" % str(1000 - 1000 + i)] + ["%9s" % ts(v) for v in vals]) + "\n")
    data = open(tmp_path / "gel1bodycenter.dat").read()
    assert data == "".join(rows)
    assert np.abs(ring["center_rm"][1:7]).min() > 0 and np.all(ring["center_rm"][7:] == 0)
