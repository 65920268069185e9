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
