"""z-slab decomposition: a chain of slabs (one GPU's share each, halo planes
exchanged through packed face buffers) reproduces the whole lattice's
stream/collide sub-steps bit for bit, with walls and periodic in z, in FP32
and FP16 storage, with uneven slab heights and a one-plane slab.

Reference: the whole lattice's own Fluid.sub_step with the marker force off
(beta_eff = 0).  Its gel flags blocks for the forced collision, which with
a zero force stores the same values as the slabs' unforced collision (in
FP16 too: both round the FP32 result to FP16).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def lattice(G, box, per, storage, seed):
    g = G.Gel((4, 4, 4), (3.0, 3.0, 3.0))
    c = G.Coupler([g])
    f = G.Fluid(box, c)
    f.set_periodic(per[0], per[1], per[2])
    f.set_storage(storage)
    rng = np.random.default_rng(seed)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    f.write("f", (w[:, None] * (1 + 0.02 * rng.standard_normal((19, f.N)))).astype(np.float32))
    return g, c, f


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("per", [(0, 0, 0), (1, 0, 1), (0, 1, 1)])
@pytest.mark.parametrize("split", ["even", "uneven"])
def test_slab_chain_equals_whole_lattice(glsm, storage, per, split):
    G = glsm
    steps = 6
    box = (7, 6, 9)  # 15 x 13 x 19 nodes
    objs = [lattice(G, box, per, storage, 3), lattice(G, box, per, storage, 3)]
    (g0, c0, whole), (g1, c1, split_f) = objs
    Nz = whole.grid[2]
    parts = [(0, 5), (5, 5), (10, 4), (14, 5)] if split == "even" else [(0, 1), (1, 7), (8, 2), (10, 9)]
    assert sum(n for _, n in parts) == Nz
    slabs = [G.Slab(split_f, z0, nz) for z0, nz in parts]
    try:
        for s in slabs:
            s.load(split_f)
        for _ in range(steps):
            whole.sub_step(0.0)
        G.slab_chain_step(slabs, steps)
        for s in slabs:
            s.store(split_f)
        np.testing.assert_array_equal(split_f.read("fpost"), whole.read("fpost"))
    finally:
        for s in slabs:
            s.close()
        for g, c, f in objs:
            f.close()
            c.close()
            g.close()


def test_single_periodic_slab_and_errors(glsm):
    """One slab covering a z-periodic lattice exchanges with itself; chains
    must be contiguous and cover the lattice."""
    G = glsm
    objs = [lattice(G, (5, 5, 5), (0, 0, 1), "f32", 5) for _ in range(2)]
    (g0, c0, whole), (g1, c1, other) = objs
    s = G.Slab(other, 0, whole.grid[2])
    a, b = G.Slab(other, 0, 3), G.Slab(other, 4, 7)
    try:
        s.load(other)
        for _ in range(4):
            whole.sub_step(0.0)
        G.slab_chain_step([s], 4)
        s.store(other)
        np.testing.assert_array_equal(other.read("fpost"), whole.read("fpost"))
        assert s.face_bytes() == 5 * whole.grid[0] * whole.grid[1] * 4
        with pytest.raises(G.GlsmError, match="contiguous"):
            G.slab_chain_step([a, b], 1)
        with pytest.raises(G.GlsmError, match="cover"):
            G.slab_chain_step([a], 1)
        with pytest.raises(G.GlsmError, match="outside"):
            G.Slab(other, 8, 5)
    finally:
        for x in (s, a, b):
            x.close()
        for g, c, f in objs:
            f.close()
            c.close()
            g.close()
