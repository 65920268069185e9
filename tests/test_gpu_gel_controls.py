"""Gel parameter and state controls against the FP64 restatement.

  * set_chemistry_rates (f, ep of GelParams, gel.cu:633-652; steady state
    solved again, gel.cu:424-443, interior chemistry reset to it,
    gel.cu:179-219) -- the batching workload's +-20% rates -- then 16
    perturbed gel steps: every field within 1e-12 of the oracle;
  * resetToQuiescent (gel.cu:731-747) in the middle of a run, as sim_cilia
    and sim_single use it: within 1e-12 after 11 more steps;
  * node positions written from the host (glsm_gel_write of rn) move the
    IBM markers at once: the next sub-step interpolates at the new
    positions, bit for bit with the FP32 restatement fed the same velocity
    field, for a single lattice and for a batch member.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_gel_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rates", [(0.72, 0.33), (1.08, 0.24)])
def test_set_chemistry_rates_vs_fp64(glsm, oracle, rates):
    g, og = glsm.Gel((6, 6, 6), (10, 10, 10)), oracle.Gel((6, 6, 6), (10, 10, 10))
    for gel in (g, og):
        gel.set_chemistry_rates(*rates)
    uss = og.steady_state()
    um = g.read("um").reshape(8, 8, 8)
    assert abs(um[3, 3, 3] - uss[0]) <= 1e-12 * uss[0]
    assert_gel_close(g, og, fields=["um", "vm", "wm"])
    for gel in (g, og):
        gel.perturb_chemistry(3, 0.01)
    for it in range(16):
        for gel in (g, og):
            gel.step_elasticity(it)
            gel.step_chemistry(it)
    assert_gel_close(g, og)
    g.close()


def test_reset_to_quiescent_vs_fp64(glsm, oracle):
    g, og = glsm.Gel((6, 4, 5), (10, 10, 10)), oracle.Gel((6, 4, 5), (10, 10, 10))
    for gel in (g, og):
        gel.perturb_chemistry(5, 0.02)
    for it in range(7):
        for gel in (g, og):
            gel.step_elasticity(it)
            gel.step_chemistry(it)
    for gel in (g, og):
        gel.reset_to_quiescent()
    assert_gel_close(g, og, fields=["um", "vm", "wm"])
    for it in range(7, 18):
        for gel in (g, og):
            gel.step_elasticity(it)
            gel.step_chemistry(it)
    assert_gel_close(g, og)
    g.close()


def coupled(mod, prec=None):
    g = mod.Gel((4, 4, 4), (6.0, 6.0, 6.0))
    c = mod.Coupler([g]) if prec is None else mod.Coupler([g], prec)
    f = mod.Fluid((12, 12, 12), c)
    return g, c, f


def velocity_field(N, seed):
    rng = np.random.default_rng(seed)
    return (1e-3 * rng.standard_normal((N, 3))).astype(np.float32)


def test_written_node_positions_move_the_markers(glsm, oracle):
    G = glsm
    g, c, f = coupled(G)
    og, oc, of = coupled(oracle, 32)
    try:
        for gel in (g, og):
            gel.step_elasticity(0)
        for cp in (c, oc):
            cp.pack_from_gels()
        for fl in (f, of):
            fl.sub_step(0.5)  # stencil tables and spreading lists built at the first positions
        rng = np.random.default_rng(8)
        rn = og.read("rn") + 0.3 * rng.standard_normal(og.read("rn").shape)
        g.write("rn", rn)
        og.write("rn", rn)
        u = velocity_field(f.N, 2)
        f.write("u", u)
        of.write("u", u.astype(np.float64))
        for cp in (c, oc):
            cp.pack_from_gels()
        np.testing.assert_array_equal(c.read("lag").astype(np.float64), oc.read("lag"))
        for fl in (f, of):
            fl.sub_step(0.5)
        np.testing.assert_array_equal(c.read("Ul").astype(np.float64), oc.read("Ul"))
        np.testing.assert_array_equal(c.read("Fdrag").astype(np.float64), oc.read("Fdrag"))
    finally:
        for o in (g, f, c):
            o.close()


def test_written_node_positions_move_a_batch_members_markers(glsm):
    """The same write between two steps of a batch member: the member's
    next step equals the instance stepped alone (which re-bins, as the
    single-lattice test above shows) bit for bit."""
    G = glsm
    runs = []
    for batched in (False, True):
        g, c, f = coupled(G)
        f.set_ibm(1.0, 20.0)
        b = G.Batch("f32") if batched else None
        try:
            if b:
                b.add([g], c, f, iteration=0)
                b.step(2)
            else:
                for it in range(2):
                    G.system_step([g], c, f, it)
            rng = np.random.default_rng(8)
            g.write("rn", g.read("rn") + 0.3 * rng.standard_normal(g.read("rn").shape))
            if b:
                b.step(1)
            else:
                G.system_step([g], c, f, 2)
            runs.append({"lag": c.read("lag"), "Ul": c.read("Ul"), "fpost": f.read("fpost"), "rn": g.read("rn")})
        finally:
            if b:
                b.close()
            for o in (g, f, c):
                o.close()
    for k in runs[0]:
        np.testing.assert_array_equal(runs[0][k], runs[1][k], err_msg=k)
