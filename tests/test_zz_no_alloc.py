"""The allocation-free conservation entry point (glsm_fluid_conservation_static,
static device scratch) against glsm_fluid_conservation (library-allocated
scratch), called through the C ABI: the same bits, on two calls in a row
with the second on a smaller lattice, in FP32 and FP16 storage; a lattice
beyond the scratch's 33,554,432 cells is rejected with GLSM_EINVAL."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)


def fluid(G, box, storage, seed):
    g = G.Gel((2, 2, 2), (box[0] / 2, box[1] / 2, box[2] / 2))
    c = G.Coupler([g])
    f = G.Fluid(box, c)
    f.set_storage(storage)
    rng = np.random.default_rng(seed)
    f.write("f", (W[:, None] + 1e-3 * rng.standard_normal((19, f.N))).astype(np.float32))
    return g, c, f


def both(G, f):
    L = G.load()
    a, b = (C.c_double * 4)(), (C.c_double * 4)()
    assert L.glsm_fluid_conservation(f.h, a) == 0
    assert L.glsm_fluid_conservation_static(f.h, b) == 0, L.glsm_last_error()
    return np.array(a[:]), np.array(b[:])


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_static_scratch_conservation_equals_allocating_one(glsm, storage):
    G = glsm
    objs = []
    try:
        for box, seed in (((60, 50, 40), 1), ((12, 9, 7), 2)):  # the second call on a smaller lattice
            g, c, f = fluid(G, box, storage, seed)
            objs += [g, f, c]
            a, b = both(G, f)
            print(storage, box, a, flush=True)
            assert np.array_equal(a, b)
            assert np.abs(a[1:]).max() > 0
            G.system_step([g], c, f, 0)
            a, b = both(G, f)
            assert np.array_equal(a, b)
    finally:
        for o in objs:
            o.close()


def test_static_scratch_rejects_a_larger_lattice(glsm):
    G = glsm
    c = G.Coupler([])
    f = G.Fluid((162, 161, 161), c)  # 325 x 323 x 323 = 33,909,125 cells
    try:
        assert f.N > 33554432
        out = (C.c_double * 4)()
        L = G.load()
        assert L.glsm_fluid_conservation_static(f.h, out) == -1
        assert b"33554432" in L.glsm_last_error()
        assert L.glsm_fluid_conservation(f.h, out) == 0
    finally:
        f.close()
        c.close()
