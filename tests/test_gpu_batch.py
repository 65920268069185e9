"""Batched stepping: every member of a Batch equals the same instance stepped
alone with system_step (sim.cpp's loop body), bit for bit -- all gel fields,
u, c and the stored post-collision populations -- for lattices and gels of
different sizes, walls and periodic boxes, varied rate constants, graph
capture on and off, FP16 storage, every stream/collide launch shape, and
members that join and leave at different iterations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (box, periodic, [(gel size, position, rates or None)], perturbation seed)
SPECS = [
    ((10, 10, 10), 1, [((4, 4, 4), (5.0, 5.0, 5.0), None)], 1),
    ((16, 16, 16), 0, [((4, 4, 4), (5.0, 8.0, 8.0), (1.0, 0.33)), ((5, 4, 5), (11.0, 8.2, 7.9), None)], 2),
    ((12, 10, 14), 0, [((6, 6, 6), (6.0, 5.0, 7.0), (0.8, 0.27))], 3),
]
FIELDS = ["rn", "um", "vm", "wm", "um_norm", "vm_norm", "wmp", "T0", "T1", "T2", "vol", "pm", "rm", "Veln",
          "Vels", "Fn", "Fn_robin", "Fdrag_robin", "un_norm", "un_robin", "vn_norm", "center_vm", "center_rm"]


def build(G, spec, storage="f32"):
    box, per, gel_specs, seed = spec
    gels = []
    for k, (size, pos, rates) in enumerate(gel_specs):
        g = G.Gel(size, pos, gel_id=k + 1)
        if rates:
            g.set_chemistry_rates(*rates)
        g.perturb_chemistry(seed * 10 + k, 0.01)
        gels.append(g)
    c = G.Coupler(gels)
    f = G.Fluid(box, c)
    f.set_periodic(per, per, per)
    f.set_ibm(1.0, 20.0)
    f.set_storage(storage)
    return gels, c, f


def state(inst):
    gels, c, f = inst
    out = {"u": f.read("u"), "c": f.read("c"), "fpost": f.read("fpost")}
    for k, g in enumerate(gels):
        for name in FIELDS:
            out[f"gel{k}.{name}"] = g.read(name)
    return out


def close(inst):
    gels, c, f = inst
    for g in gels:
        g.close()
    f.close()
    c.close()


def assert_same(a, b, what):
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


def solo(G, spec, start, steps, storage="f32"):
    inst = build(G, spec, storage)
    for it in range(start, start + steps):
        G.system_step(inst[0], inst[1], inst[2], it)
    st = state(inst)
    close(inst)
    return st


@pytest.mark.parametrize("graph", [True, False])
def test_batch_members_equal_solo_runs(glsm, graph):
    G = glsm
    steps = 12
    ref = [solo(G, s, 0, steps) for s in SPECS]
    b = G.Batch("f32")
    insts = [build(G, s) for s in SPECS]
    try:
        b.use_graph(graph)
        ids = [b.add(*inst) for inst in insts]
        b.step(steps)
        for mid in ids:
            assert b.iteration(mid) == steps
        for k, inst in enumerate(insts):
            assert_same(state(inst), ref[k], f"member {k}")
    finally:
        b.close()
        for inst in insts:
            close(inst)


def test_batch_fp16_and_every_launch_shape(glsm):
    G = glsm
    ref = [solo(G, s, 0, 6, "f16") for s in SPECS]
    shapes = [(nt, run, grid, 1, 1) for nt in (0, 1) for run in (0, 4, 16, 64) for grid in (0, 3, 256)]
    # larger work-groups and several rounds of blocks per work-group; units
    # of up to 8 blocks straddle the members' block ranges
    shapes += [(nt, run, 0, wgb, chunk) for nt in (0, 1) for run in (0, 4, 16)
               for wgb, chunk in ((2, 1), (4, 1), (1, 3), (1, 8), (2, 2))]
    # LDS reserved per work-group (fewer work-groups resident per CU)
    shapes += [(nt, run, 0, 1, 1, kb * 1024) for nt in (0, 1) for run in (4, 16) for kb in (32, 52, 64)]
    for shape in shapes:
        b = G.Batch("f16")
        insts = [build(G, s, "f16") for s in SPECS]
        try:
            for inst in insts:
                b.add(*inst)
            b.set_launch(*shape)
            b.step(6)
            for k, inst in enumerate(insts):
                assert_same(state(inst), ref[k], f"shape {shape} member {k}")
        finally:
            b.close()
            for inst in insts:
                close(inst)


def test_batch_autotune_keeps_state_and_results(glsm):
    G = glsm
    ref = [solo(G, s, 0, 8) for s in SPECS]
    b = G.Batch("f32")
    insts = [build(G, s) for s in SPECS]
    try:
        for inst in insts:
            b.add(*inst)
        b.step(3)
        before = [state(inst) for inst in insts]
        shape, ms = b.autotune(reps=2)
        assert ms > 0 and shape["run"] in (0, 4, 16, 64) and shape["wgb"] in (1, 2, 4) and 1 <= shape["chunk"] <= 8
        assert shape["lds"] in (0, 32 * 1024, 40 * 1024, 52 * 1024)
        for k, inst in enumerate(insts):
            assert_same(state(inst), before[k], f"autotune changed member {k}")
        b.step(5)
        for k, inst in enumerate(insts):
            assert_same(state(inst), ref[k], f"member {k}")
    finally:
        b.close()
        for inst in insts:
            close(inst)


def test_batch_members_join_and_leave(glsm):
    """Member 0 runs 0..11; member 1 joins after 4 batch steps at its own
    iteration 100 and runs 100..107; member 2 runs 0..5 and leaves, then is
    stepped alone 6..9 with system_step."""
    G = glsm
    ref0 = solo(G, SPECS[0], 0, 12)
    ref1 = solo(G, SPECS[1], 100, 8)
    ref2 = solo(G, SPECS[2], 0, 10)
    b = G.Batch("f32")
    i0, i1, i2 = build(G, SPECS[0]), build(G, SPECS[1]), build(G, SPECS[2])
    try:
        m0 = b.add(*i0)
        m2 = b.add(*i2)
        b.step(4)
        m1 = b.add(*i1, iteration=100)
        b.step(2)
        assert b.iteration(m2) == 6
        b.remove(m2)
        for it in range(6, 10):
            G.system_step(i2[0], i2[1], i2[2], it)
        b.step(6)
        assert b.iteration(m0) == 12 and b.iteration(m1) == 108
        assert_same(state(i0), ref0, "member 0")
        assert_same(state(i1), ref1, "member 1")
        assert_same(state(i2), ref2, "member 2 (left the batch)")
    finally:
        b.close()
        for inst in (i0, i1, i2):
            close(inst)


def test_member_stepped_outside_the_batch(glsm):
    """A member advanced by its own methods between batch steps (here one
    extra LBM sub-step, which swaps its buffers) is picked up from its
    current buffers, as in the same sequence run alone."""
    G = glsm
    spec = SPECS[1]
    inst = build(G, spec)
    for it in range(3):
        G.system_step(inst[0], inst[1], inst[2], it)
    inst[2].sub_step(0.5)
    for it in range(3, 6):
        G.system_step(inst[0], inst[1], inst[2], it)
    ref = state(inst)
    close(inst)
    b = G.Batch("f32")
    inst = build(G, spec)
    try:
        b.add(*inst)
        b.step(3)
        inst[2].sub_step(0.5)
        b.step(3)
        assert_same(state(inst), ref, "member stepped outside")
    finally:
        b.close()
        close(inst)


def test_batch_rejects_bad_members(glsm):
    G = glsm
    b = G.Batch("f32")
    inst = build(G, SPECS[0])
    try:
        b.add(*inst)
        with pytest.raises(G.GlsmError, match="already a member"):
            b.add(*inst)
        with pytest.raises(G.GlsmError, match="no member"):
            b.remove(12345)
        with pytest.raises(G.GlsmError, match="run must be 0, 4, 16 or 64"):
            b.set_launch(0, 3, 0)
        with pytest.raises(G.GlsmError, match="several blocks"):
            b.set_launch(0, 4, 256, 2, 1)
        with pytest.raises(G.GlsmError, match="lds must be in"):
            b.set_launch(0, 4, 0, 1, 1, 70000)
        with pytest.raises(G.GlsmError, match="LDS reserve takes one block"):
            b.set_launch(0, 4, 0, 2, 1, 32768)
    finally:
        b.close()
        close(inst)


@pytest.mark.parametrize("mode,name", [(2, "thread form"), (3, "staged form")])
def test_interpolation_forms_identical(glsm, mode, name):
    """One wavefront per marker (solo runs), one thread per marker and the
    LDS-staged form (batch and solo) sum the same 64 products in the same
    order: identical bits."""
    G = glsm
    try:
        G.set_interp_mode(1)
        ref = [solo(G, s, 0, 8) for s in SPECS]
        G.set_interp_mode(mode)
        b = G.Batch("f32")
        insts = [build(G, s) for s in SPECS]
        try:
            for inst in insts:
                b.add(*inst)
            b.step(8)
            for k, inst in enumerate(insts):
                assert_same(state(inst), ref[k], f"{name}, member {k}")
        finally:
            b.close()
            for inst in insts:
                close(inst)
        assert_same(solo(G, SPECS[1], 0, 8), ref[1], f"{name}, solo")
    finally:
        G.set_interp_mode(0)


def _per_object_step(gels, c, f, it):
    """sim.cpp's loop body through the objects' own methods, one gel at a time."""
    for g in gels:
        g.step_elasticity(it)
        g.step_chemistry(it)
        g.record_center(it)
    c.pack_from_gels()
    f.step_velocity(it)
    f.step_concentration(it)
    c.apply_gel_repulsion()
    c.scatter_to_gels()


def _change_params(inst, k):
    gels, c, f = inst
    if k == 4:
        gels[-1].set_chemistry_rates(0.9, 0.3)
    if k == 7:
        gels[0].set_anchor_z(2.5)
        f.set_ibm(0.8, 10.0)


def test_system_step_equals_per_object_steps_with_changes(glsm):
    """glsm_system_step launches each gel phase once for all the coupler's
    gels; it equals the objects' own per-gel calls bit for bit, also when
    gel and IBM parameters change between steps."""
    G = glsm
    spec = SPECS[1]  # two gels
    a, b = build(G, spec), build(G, spec)
    try:
        for it in range(11):
            _change_params(a, it)
            _change_params(b, it)
            G.system_step(a[0], a[1], a[2], it)
            _per_object_step(*b, it)
        assert_same(state(a), state(b), "system step vs per-object steps")
    finally:
        close(a)
        close(b)


def test_batch_sees_parameter_changes(glsm):
    """A member's gel and IBM parameters changed through its own objects
    between batch steps are used from the next step on, as in its solo run."""
    G = glsm
    spec = SPECS[1]
    solo_inst, member = build(G, spec), build(G, spec)
    b = G.Batch("f32")
    try:
        b.add(*member)
        for it in range(11):
            _change_params(solo_inst, it)
            _change_params(member, it)
            G.system_step(solo_inst[0], solo_inst[1], solo_inst[2], it)
            b.step(1)
        assert_same(state(member), state(solo_inst), "batch member with parameter changes")
    finally:
        b.close()
        close(solo_inst)
        close(member)


def test_autotune_cache_reuses_the_pick(glsm, tmp_path):
    """A pick cached under the batch's workload key is set again without
    timing; another workload gets its own entry."""
    import json

    G = glsm
    path = str(tmp_path / "shapes.json")
    insts = [build(G, s) for s in SPECS]
    try:
        b = G.Batch("f32")
        for inst in insts:
            b.add(*inst)
        shape, ms = b.autotune(reps=1, cache=path)
        data = json.load(open(path))
        assert list(data) == [b.workload_key()] and data[b.workload_key()]["launch"] == shape
        b.set_launch(0, 64, 256, 1, 1)
        again, ms2 = b.autotune(reps=1, cache=path)
        assert again == shape and ms2 == ms
        b.close()
        b = G.Batch("f32")
        b.add(*insts[0])
        b.autotune(reps=1, cache=path)
        assert len(json.load(open(path))) == 2
        b.close()
    finally:
        for inst in insts:
            close(inst)
