"""Parity of the HIP path (through the C ABI) with the CPU restatement.

Tolerances:
  * topology / index maps: bit-exact;
  * gel (FP64 on both sides): relative 1e-12 per field (the device log/sqrt
    may differ from the host's in the last bit);
  * IBM interpolation, marker forces, coupler packing, repulsion, the
    concentration step: bit-exact against the FP32 restatement when fed the
    same inputs (both evaluate the reference's operation order unfused);
  * one LBM sub-step: FP32 round-off (rel-L2 <= 3e-7 on f) against the FP32
    restatement for the reference's full-f storage, and against FP64 for the
    default deviation storage; FP16 / BF16 deviation storage within
    2e-4 / 2e-3 on u for one sub-step;
  * the small plain-method workload (41^3, one gel, periodic) against FP64:
    velocity rel-L2 <= 1e-5, concentration rel-L2 <= 1e-6, gel node
    positions within 5e-7 gel units (= 1e-6 lattice units) after 200 live
    steps and after 1000 steps against tests/golden/cube41_one_gel.npz.
"""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, 1], [-1, 0, -1],
              [1, 0, -1], [0, 1, 1], [0, 1, -1], [0, -1, -1], [0, -1, 1]])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def feq(rho, u):
    cu = u @ C.T
    uu = (u * u).sum(1)[:, None]
    return (W[None] * rho[:, None] * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * uu)).T


GEL_FIELDS_CMP = ["rn", "um", "vm", "wm", "um_norm", "vm_norm", "wmp", "T0", "T1", "T2", "vol", "pm", "rm",
                  "rm_loc", "nmSm", "Veln", "Vels", "Fn", "un_norm", "vn_norm", "un_robin"]


def assert_gel_close(g, og, tol=1e-12, fields=GEL_FIELDS_CMP):
    """Each field within `tol` of the oracle, relative to the field's largest
    magnitude.  T1 = (sum of the six face area vectors) . vmum / vol vanishes
    identically for a closed hexahedron (the 2x2 rule is exact for bilinear
    faces), so both sides hold rounding noise; it is compared on the scale of
    its terms, |nmSm| |Veln| |un_norm| / vol."""
    bad = []
    for name in fields:
        a, b = g.read(name), og.read(name)
        if name == "T1":
            scale = 8 * np.abs(og.read("nmSm")).max() * np.abs(og.read("Veln")).max() * \
                np.abs(og.read("un_norm")).max() / np.abs(og.read("vol")).max()
        else:
            scale = np.abs(b).max()
        err = np.abs(a - b).max() / max(scale, 1e-300)
        if err > tol:
            bad.append(f"{name}: {err:.3e}")
    assert not bad, "max relative errors: " + ", ".join(bad)


# ------------------------------------------------------------------ topology
@pytest.mark.parametrize("size", [(1, 1, 1), (6, 6, 6), (8, 8, 8), (4, 4, 24), (20, 20, 4), (3, 5, 2)])
def test_topology_bit_exact(glsm, oracle, size):
    say("topology", size)
    g, og = glsm.Gel(size, (40, 40, 40)), oracle.Gel(size, (40, 40, 40))
    assert g.node_grid == og.node_grid and g.boundary_count == og.boundary_count
    for name in ("map_node", "map_element", "bIndex"):
        assert np.array_equal(g.read(name), og.read(name)), name
    assert g.get_index(2, 3, 1, 2) == 2 + 3 * (g.node_grid[0] + 2) + 1 * (g.node_grid[0] + 2) * (g.node_grid[1] + 2)
    g.close()


# ------------------------------------------------------------------ gel (FP64)
GEL_CASES = {
    "perturbed_deformed": dict(size=(6, 6, 6), pos=(10, 10, 10), setup="deform"),
    "anchor_L8": dict(size=(7, 3, 3), pos=(10, 10, 10), setup="anchor"),
    "tube_square": dict(size=(10, 6, 6), pos=(10, 10, 10), setup="tube"),
    "tube_cylinder": dict(size=(10, 8, 8), pos=(10, 10, 10), setup="cyl"),
    "type2_pulse": dict(size=(5, 4, 3), pos=(10, 10, 10), setup="type2"),
}


@pytest.mark.parametrize("case", sorted(GEL_CASES))
def test_gel_steps_match_fp64(glsm, oracle, case):
    say("gel case", case)
    cfg = GEL_CASES[case]
    gtype = 2 if cfg["setup"] == "type2" else 1
    g = glsm.Gel(cfg["size"], cfg["pos"], gel_type=gtype)
    og = oracle.Gel(cfg["size"], cfg["pos"], gel_type=gtype)
    for gel in (g, og):
        gel.perturb_chemistry(7, 0.02)
        if cfg["setup"] == "anchor":
            gel.set_anchor_z(7.5)
        elif cfg["setup"] == "tube":
            gel.build_tube_mask(2, False)
        elif cfg["setup"] == "cyl":
            gel.build_tube_mask(2, True)
        elif cfg["setup"] == "type2":
            gel.fire_excitation_pulse()
    if cfg["setup"] == "deform":
        rng = np.random.default_rng(11)
        rn = og.read("rn")
        rn = rn + 0.03 * rng.standard_normal(rn.shape)
        g.write("rn", rn)
        og.write("rn", rn)
    for it in range(16):
        for gel in (g, og):
            gel.step_elasticity(it)
            gel.step_chemistry(it)
            gel.record_center(it)
    assert_gel_close(g, og)
    for name in ("center_vm", "center_wm", "center_rm", "center_Fn", "center_Veln", "filament"):
        a, b = g.read(name), og.read(name)
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), name
    if cfg["setup"] == "anchor":
        L = g.node_grid
        rn = g.read("rn").reshape(L[2] + 2, L[1] + 2, L[0] + 2, 3)
        assert np.all(rn[1, 1:L[1] + 1, 1:L[0], 2] == 7.5)  # x = L (=8) is outside the anchor launch
        assert rn[1, 1, L[0], 2] != 7.5
    g.close()


# ------------------------------------------------------------------ coupler
def two_gel_system(glsm, oracle, prec=32, per=0, box=(16, 12, 12), storage="f32"):
    specs = [((4, 4, 4), (5.0, 6.0, 6.0)), ((3, 4, 3), (9.6, 6.2, 5.9))]  # overlapping repulsion range
    gs = [glsm.Gel(s, p, gel_id=i + 1) for i, (s, p) in enumerate(specs)]
    ogs = [oracle.Gel(s, p, gel_id=i + 1) for i, (s, p) in enumerate(specs)]
    for i, (a, b) in enumerate(zip(gs, ogs)):
        a.perturb_chemistry(100 + i, 0.01)
        b.perturb_chemistry(100 + i, 0.01)
    c, oc = glsm.Coupler(gs), oracle.Coupler(ogs, prec)
    f, of = glsm.Fluid(box, c), oracle.Fluid(box, oc)
    f.set_storage(storage)
    for fl in (f, of):
        fl.set_periodic(per, per, per)
    return gs, c, f, ogs, oc, of


@pytest.mark.parametrize("cull", [4096, 1])
def test_pack_repulsion_scatter(glsm, oracle, cull):
    say(f"coupler pack/repulsion/scatter (far gels skipped from {cull} markers)")
    gs, c, f, ogs, oc, of = two_gel_system(glsm, oracle)
    glsm.set_repulsion_cull(cull)
    try:
        for it in range(2):
            glsm.system_step(gs, c, f, it)
            oracle.system_step(ogs, oc, of, it)
    finally:
        glsm.set_repulsion_cull(4096)
    for name in ("lag", "Vl", "Cl"):
        assert np.array_equal(c.read(name).astype(np.float64), oc.read(name)), name
    Fr, oFr = c.read("Frep").astype(np.float64), oc.read("Frep")
    assert np.abs(oFr).max() > 0  # the two gels are within the cut-off
    assert np.array_equal(Fr, oFr)
    assert np.array_equal(c.read("owner"), oc.read("owner").astype(np.int32))
    for g, og in zip(gs, ogs):
        assert rel_l2(g.read("un_robin"), og.read("un_robin")) < 1e-6
        assert np.array_equal(g.read("Fn_robin"), og.read("Fn_robin")) or \
            np.abs(g.read("Fn_robin") - og.read("Fn_robin")).max() < 1e-12


# ------------------------------------------------------------------ one LBM sub-step
@pytest.mark.parametrize("per", [0, 1])
@pytest.mark.parametrize("storage", ["f32full", "f32", "f16", "bf16"])
def test_lbm_substep(glsm, oracle, storage, per):
    substep_case(glsm, oracle, storage, per, (10, 10, 10))


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("per,pos", [(0, (3.4, 10.0, 16.6)), (1, (1.0, 10.0, 19.5))],
                         ids=["clipped-at-walls", "wrapped-periodic"])
def test_lbm_substep_stencils_at_the_boundary(glsm, oracle, storage, per, pos):
    """Marker stencils cut by the walls (clipped to the lattice, weights
    renormalised by the clipped sum) and, in a periodic box, marker stencils
    that wrap across the faces: the same bits as the FP32 restatement for Ul
    and Fdrag, the sub-step within the storage's tolerance of FP64."""
    g, grid = substep_case(glsm, oracle, storage, per, pos)
    lo, hi = np.floor(g).min(axis=0) - 1, np.floor(g).max(axis=0) + 2  # stencil extent per axis
    outside = (lo < 0) | (hi > np.array(grid) - 1)
    assert outside[0] and outside[2], (lo, hi, grid)  # some stencils leave the lattice in x and z


def substep_case(glsm, oracle, storage, per, pos):
    say("sub-step", storage, per, pos)
    g, og, og64 = (x.Gel((6, 6, 6), pos) for x in (glsm, oracle, oracle))
    c, oc, oc64 = glsm.Coupler([g]), oracle.Coupler([og], 32), oracle.Coupler([og64], 64)
    f, of, of64 = glsm.Fluid((20, 20, 20), c), oracle.Fluid((20, 20, 20), oc), oracle.Fluid((20, 20, 20), oc64)
    for fl in (f, of, of64):
        fl.set_periodic(per, per, per)
    f.set_storage(storage)
    Nx, Ny, Nz = f.grid
    z, y, x = np.meshgrid(np.arange(Nz), np.arange(Ny), np.arange(Nx), indexing="ij")
    u0 = np.stack([0.02 * np.sin(2 * np.pi * y / Ny), 0.015 * np.cos(2 * np.pi * z / Nz),
                   0.01 * np.sin(2 * np.pi * x / Nx)], -1).reshape(f.N, 3)
    rho0 = 1 + 0.001 * np.cos(2 * np.pi * (x + y) / Nx).reshape(f.N)
    rng = np.random.default_rng(1)
    f0 = (feq(rho0, u0) + 1e-4 * rng.standard_normal((19, f.N))).astype(np.float32)
    u0 = u0.astype(np.float32)
    for gel in (g, og, og64):
        gel.step_elasticity(0)
    for cp in (c, oc, oc64):
        cp.pack_from_gels()
    f.write("f", f0)
    f.write("u", u0)
    for fl in (of, of64):
        fl.write("f", f0.astype(np.float64))
        fl.write("u", u0.astype(np.float64))
    if storage in ("f32full", "f32"):
        assert np.array_equal(f.read("f"), f0)  # storage round trip (pull o push = id)
    for fl in (f, of, of64):
        fl.sub_step(0.5)
    # marker interpolation and forces: bit-exact with the FP32 restatement
    assert np.array_equal(c.read("Ul").astype(np.float64), oc.read("Ul"))
    assert np.array_equal(c.read("Fdrag").astype(np.float64), oc.read("Fdrag"))
    fg, ug = f.read("f"), f.read("u")
    ref_f, ref_u = (of.read("f"), of.read("u")) if storage == "f32full" else (of64.read("f"), of64.read("u"))
    tol_f = {"f32full": 3e-7, "f32": 1e-7, "f16": 2e-5, "bf16": 1e-4}[storage]
    tol_u = {"f32full": 1e-6, "f32": 2e-7, "f16": 2e-4, "bf16": 2e-3}[storage]
    assert rel_l2(fg, ref_f) < tol_f
    assert rel_l2(ug, ref_u) < tol_u
    return c.read("lag").astype(np.float64) / 0.5, f.grid  # marker positions in lattice units


# ------------------------------------------------------------------ concentration
def test_concentration_step_bit_exact(glsm, oracle):
    """The scalar step (sampling, Robin source spreading, upwind advection,
    diffusion) with walls and periodic faces, for a central gel and for gels
    whose stencils are clipped by the walls or wrap across the faces."""
    say("concentration step")
    for per, pos in ((0, (10, 10, 10)), (1, (10, 10, 10)), (0, (3.4, 10.0, 16.6)), (1, (1.0, 10.0, 19.5))):
        g, og = glsm.Gel((6, 6, 6), pos), oracle.Gel((6, 6, 6), pos)
        c, oc = glsm.Coupler([g]), oracle.Coupler([og], 32)
        f, of = glsm.Fluid((20, 20, 20), c), oracle.Fluid((20, 20, 20), oc)
        for fl in (f, of):
            fl.set_periodic(per, per, per)
        rng = np.random.default_rng(3)
        u = (1e-2 * rng.standard_normal((f.N, 3))).astype(np.float32)
        cc = rng.uniform(0, 0.2, f.N).astype(np.float32)
        f.write("u", u)
        f.write("c", cc)
        of.write("u", u.astype(np.float64))
        of.write("c", cc.astype(np.float64))
        for gel in (g, og):
            gel.step_elasticity(0)
        c.pack_from_gels()
        oc.pack_from_gels()
        f.step_concentration(0)
        of.step_concentration(0)
        assert np.array_equal(c.read("Sl").astype(np.float64), oc.read("Sl"))
        assert np.array_equal(f.read("c").astype(np.float64), of.read("c"))


# ------------------------------------------------------------------ whole steps
def cube41_one_gel(mod, prec=None, storage=None):
    g = mod.Gel((6, 6, 6), (10, 10, 10))
    c = mod.Coupler([g]) if prec is None else mod.Coupler([g], prec)
    f = mod.Fluid((20, 20, 20), c)
    f.set_periodic(1, 1, 1)
    if storage:
        f.set_storage(storage)
    g.perturb_chemistry(0, 0.01)
    return g, c, f


def test_cube41_200_steps_vs_fp64(glsm, oracle):
    say("41^3 one gel, 200 steps")
    og, oc, of = cube41_one_gel(oracle, 64)
    sims = {st: cube41_one_gel(glsm, storage=st) for st in ("f32", "f16", "bf16")}
    for it in range(200):
        oracle.system_step([og], oc, of, it)
        for g, c, f in sims.values():
            glsm.system_step([g], c, f, it)
    uo, co = of.read("u"), of.read("c")
    g, c, f = sims["f32"]
    errs = {st: rel_l2(s[2].read("u"), uo) for st, s in sims.items()}
    say("u rel-L2 vs FP64 after 200 steps:", errs)
    assert errs["f32"] <= 1e-5
    assert errs["f16"] <= 1e-3
    assert errs["bf16"] <= 1e-2
    assert rel_l2(f.read("c"), co) <= 1e-6
    assert np.abs(g.read("rn") - og.read("rn")).max() <= 5e-7
    for name in ("um", "vm", "wm"):
        assert rel_l2(g.read(name), og.read(name)) <= 1e-9, name
    # un_norm carries un_robin = -Sl from the FP32 concentration exchange
    assert rel_l2(g.read("un_norm"), og.read("un_norm")) <= 1e-6


def test_cube41_1000_steps_vs_golden(glsm):
    say("41^3 one gel, 1000 steps vs golden")
    gold = np.load(os.path.join(HERE, "golden", "cube41_one_gel.npz"))
    g, c, f = cube41_one_gel(glsm, storage="f32")
    for it in range(1000):
        glsm.system_step([g], c, f, it)
        if it + 1 == 100:
            assert np.abs(g.read("rn") - gold["rn_100"]).max() <= 5e-7
            assert rel_l2(g.read("um"), gold["um_100"]) <= 1e-9
    eu = rel_l2(f.read("u"), gold["u_1000"])
    say("u rel-L2 after 1000 steps:", eu)
    assert eu <= 1e-5
    assert rel_l2(f.read("c"), gold["c_1000"]) <= 1e-6
    assert np.abs(g.read("rn") - gold["rn_1000"]).max() <= 5e-7
    for name in ("um", "vm", "wm"):
        assert rel_l2(g.read(name), gold[f"{name}_1000"]) <= 1e-9, name
    assert rel_l2(g.read("un_norm"), gold["un_norm_1000"]) <= 1e-6


def test_walled_two_gels_vs_fp64(glsm, oracle):
    say("walls, two gels")
    gs, c, f, ogs, oc, of = two_gel_system(glsm, oracle, prec=64, per=0)
    for it in range(60):
        glsm.system_step(gs, c, f, it)
        oracle.system_step(ogs, oc, of, it)
    assert rel_l2(f.read("u"), of.read("u")) <= 1e-5
    assert rel_l2(f.read("c"), of.read("c")) <= 1e-6
    for g, og in zip(gs, ogs):
        assert np.abs(g.read("rn") - og.read("rn")).max() <= 5e-7


def test_runs_are_bitwise_deterministic(glsm):
    say("determinism")
    outs = []
    for _ in range(2):
        g, c, f = cube41_one_gel(glsm)
        for it in range(30):
            glsm.system_step([g], c, f, it)
        outs.append((f.read("u"), f.read("c"), f.read("fpost"), g.read("rn"), g.read("um")))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
