"""Edge cases of the time step: an instance with no gel (no markers: no
interpolation, no flagged block, no gel launch), alone and as a batch member
beside instances with gels."""
import numpy as np
import pytest

from test_gpu_batch import SPECS, build, close, state, assert_same

pytestmark = pytest.mark.gpu


def empty_instance(G, box=(6, 5, 7), storage="f32", kick=True):
    c = G.Coupler([])
    f = G.Fluid(box, c)
    f.set_storage(storage)
    if kick:  # a non-rest state, so the step has something to move
        rng = np.random.default_rng(11)
        w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
        f.write("f", (w[:, None] * (1 + 0.01 * rng.standard_normal((19, f.N)))).astype(np.float32))
    return [], c, f


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_instance_without_gels(glsm, storage):
    G = glsm
    gels, c, f = empty_instance(G, storage=storage, kick=False)
    try:
        for it in range(3):
            G.system_step(gels, c, f, it)
        # a rest state with no force stays at rest, its mass exactly conserved
        assert np.array_equal(f.read("u"), np.zeros_like(f.read("u")))
        cons = f.conservation()
        assert cons[0] == 0.0 and np.all(cons[1:] == 0.0)
    finally:
        f.close()
        c.close()


def test_batch_member_without_gels(glsm):
    G = glsm
    solo = empty_instance(G)
    for it in range(6):
        G.system_step(solo[0], solo[1], solo[2], it)
    ref = {"u": solo[2].read("u"), "fpost": solo[2].read("fpost"), "c": solo[2].read("c")}
    solo[2].close()
    solo[1].close()
    member, other = empty_instance(G), build(G, SPECS[0])
    b = G.Batch("f32")
    try:
        b.add(*member)
        b.add(*other)
        b.step(6)
        got = {"u": member[2].read("u"), "fpost": member[2].read("fpost"), "c": member[2].read("c")}
        assert_same(got, ref, "gel-free member")
    finally:
        b.close()
        member[2].close()
        member[1].close()
        close(other)


def test_repulsion_skipping_far_gels_is_exact(glsm):
    """Three gels, two within the repulsion cut-off of each other and one far
    away: skipping the gels whose markers' bounding box is out of range gives
    the full scan's forces bit for bit."""
    G = glsm
    specs = [((4, 4, 4), (5.0, 6.0, 6.0)), ((3, 4, 3), (9.6, 6.2, 5.9)), ((3, 3, 3), (20.0, 18.0, 17.0))]

    def system():
        gels = [G.Gel(s, p, gel_id=i + 1) for i, (s, p) in enumerate(specs)]
        for i, g in enumerate(gels):
            g.perturb_chemistry(100 + i, 0.01)
        c = G.Coupler(gels)
        return gels, c, G.Fluid((26, 22, 22), c)

    out = []
    for cull in (0, 1):
        G.set_repulsion_cull(cull)
        try:
            gels, c, f = system()
            for it in range(3):
                G.system_step(gels, c, f, it)
            out.append({"Frep": c.read("Frep"), "u": f.read("u"),
                        **{f"g{k}.Fn_robin": g.read("Fn_robin") for k, g in enumerate(gels)}})
            assert np.abs(out[-1]["Frep"]).max() > 0  # the two near gels repel
            f.close()
            c.close()
            for g in gels:
                g.close()
        finally:
            G.set_repulsion_cull(4096)
    assert_same(out[1], out[0], "culled vs full repulsion")


@pytest.mark.parametrize("src,dst", [("f32", "f16"), ("f16", "bf16"), ("bf16", "f32full"), ("f32full", "f32")])
def test_storage_conversion_equals_a_fresh_import(glsm, src, dst):
    """setStorage converts the stored state on the device: the result equals
    writing the same pre-collision populations into a fluid that already
    has the target storage, bit for bit; converting to the current storage
    changes nothing."""
    G = glsm
    _, c1, f1 = empty_instance(G, storage=src)
    _, c2, f2 = empty_instance(G, storage=dst, kick=False)
    try:
        G.system_step([], c1, f1, 0)
        pops = f1.read("f")
        before = f1.read("fpost")
        f1.set_storage(src)
        assert np.array_equal(f1.read("fpost"), before)
        f1.set_storage(dst)
        f2.write("f", pops)
        assert np.array_equal(f1.read("fpost"), f2.read("fpost"))
    finally:
        for o in (f1, c1, f2, c2):
            o.close()
