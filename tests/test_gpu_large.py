"""A lattice whose FP32 distribution buffers exceed 4 GiB takes the
stream/collide kernel's 64-bit addressing (smaller ones address by 32-bit
byte offsets).  Check it against the 32-bit path through periodic tiling:
a periodic 429 x 429 x 351 lattice initialised with 11 x 11 x 9 copies of a
periodic 39^3 state must, after one sub-step, hold the same copies of the
39^3 lattice's result, bit for bit (each cell's update reads only its
neighbours).  The IBM force is off (beta_eff = 0), so the gel markers the
API requires add exactly zero.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TILE = 39             # small lattice edge (grid = 2 * size + 1)
REPS = (9, 11, 11)    # copies along z, y, x


def make(G, grid_zyx):
    g = G.Gel((4, 4, 4), (4.0, 4.0, 4.0))
    c = G.Coupler([g])
    size = tuple((n - 1) // 2 for n in reversed(grid_zyx))  # (x, y, z) box sizes
    f = G.Fluid(size, c)
    assert f.grid == tuple(reversed(grid_zyx)), (f.grid, grid_zyx)
    f.set_periodic(1, 1, 1)
    return g, c, f


def test_large_lattice_64bit_addressing_matches_tiled_small(glsm):
    G = glsm
    small_zyx = (TILE, TILE, TILE)
    big_zyx = tuple(TILE * r for r in REPS)
    rng = np.random.default_rng(7)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)
    f0 = (w[:, None] * (1 + 0.01 * rng.standard_normal((19, TILE ** 3)))).astype(np.float32)

    gs, cs, fs = make(G, small_zyx)
    gb, cb, fb = make(G, big_zyx)
    try:
        assert 19 * fb.N * 4 >= 1 << 32  # the 64-bit path
        fs.write("f", f0)
        big0 = np.tile(f0.reshape(19, *small_zyx), (1,) + REPS).reshape(19, -1)
        fb.write("f", big0)
        del big0
        fs.sub_step(0.0)
        fb.sub_step(0.0)
        G.synchronize()
        small = fs.read("f").reshape(19, *small_zyx)
        big = fb.read("f").reshape(19, *big_zyx)
        for rz in range(REPS[0]):
            for ry in range(REPS[1]):
                blk = big[:, rz * TILE:(rz + 1) * TILE, ry * TILE:(ry + 1) * TILE, :]
                np.testing.assert_array_equal(blk.reshape(19, TILE, TILE, REPS[2], TILE),
                                              np.broadcast_to(small[:, :, :, None, :], (19, TILE, TILE, REPS[2], TILE)),
                                              err_msg=f"tile z{rz} y{ry}")
    finally:
        for o in (fs, fb, cs, cb, gs, gb):
            o.close()


def test_batch_rejects_a_member_over_4gib(glsm):
    """The batched kernels address a member's buffers by 32-bit byte
    offsets: Batch.add refuses a lattice whose FP32 distributions reach
    4 GiB, and the batch stays usable."""
    G = glsm
    g, c, f = make(G, tuple(TILE * r for r in REPS))
    gs, cs, fs = make(G, (TILE, TILE, TILE))
    b = G.Batch("f32")
    try:
        assert 19 * f.N * 4 >= 1 << 32
        with pytest.raises(G.GlsmError, match="too large for a batch member"):
            b.add([g], c, f)
        mid = b.add([gs], cs, fs)
        b.step(1)
        assert b.iteration(mid) == 1
    finally:
        b.close()
        for o in (f, fs, c, cs, g, gs):
            o.close()
