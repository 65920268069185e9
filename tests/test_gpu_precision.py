"""Precision form: FP64 conservation reductions and per-field error.

  * The FP64 reduction kernel (Fluid::conservation) is pinned in the FP32
    full-f storage, where the pre-collision populations read back through the
    C ABI are exactly the values the kernel sums: its result equals a float64
    host sum of them (tolerance: FP64 summation-order round-off).
  * Against the FP64 oracle on the walled two-gel case, in every storage:
    the momentum sum within the storage's velocity tolerance on the L1 scale
    of the local momentum (FP32 1e-5, FP16 1e-3, BF16 1e-2), and the mass
    excess sum(rho - 1) within 1e-8 * N (FP32), 1e-6 * N (FP16), 2e-6 * N
    (BF16) after 1000 sub-steps (printed; BGK and Guo forcing conserve mass
    exactly, so only rounding moves it).
"""
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, 1], [-1, 0, -1],
              [1, 0, -1], [0, 1, 1], [0, 1, -1], [0, -1, -1], [0, -1, 1]], dtype=np.float64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def host_moments(f):
    """float64 (sum(rho - 1), sum(rho u)) and the L1 of the local momentum
    of a (19, N) population array."""
    f = np.asarray(f, dtype=np.float64)
    rho_m1 = f.sum(axis=0) - 1.0
    m = C.T @ f  # (3, N)
    return np.concatenate([[rho_m1.sum()], m.sum(axis=1)]), np.abs(m).sum()


def walled_two_gels(mod, prec=None, storage=None, box=(16, 12, 12)):
    specs = [((4, 4, 4), (5.0, 6.0, 6.0)), ((3, 4, 3), (9.6, 6.2, 5.9))]
    gs = [mod.Gel(s, p, gel_id=i + 1) for i, (s, p) in enumerate(specs)]
    for i, g in enumerate(gs):
        g.perturb_chemistry(100 + i, 0.01)
    c = mod.Coupler(gs) if prec is None else mod.Coupler(gs, prec)
    f = mod.Fluid(box, c)
    f.set_ibm(1.0, 20.0)  # ramp over 20 steps so the gels drive the fluid quickly
    if storage:
        f.set_storage(storage)
    return gs, c, f


def close_all(*objs):
    for o in objs:
        o.close()


def test_conservation_kernel_matches_host_fp64_sum(glsm):
    gs, c, f = walled_two_gels(glsm, storage="f32full")
    try:
        for it in range(30):
            glsm.system_step(gs, c, f, it)
        got = f.conservation()
        want, l1 = host_moments(f.read("f").reshape(19, -1))
        say("conservation f32full:", got, "host:", want, "L1:", l1)
        assert l1 > 0
        np.testing.assert_allclose(got[0], want[0], rtol=1e-9, atol=1e-12 * f.N)
        np.testing.assert_allclose(got[1:], want[1:], rtol=1e-9, atol=1e-12 * l1)
    finally:
        close_all(*gs, f, c)


@pytest.mark.parametrize("storage,tol,mass_tol", [("f32", 1e-5, 1e-8), ("f16", 1e-3, 1e-6), ("bf16", 1e-2, 2e-6)])
def test_conservation_vs_fp64_oracle(glsm, oracle, storage, tol, mass_tol):
    gs, c, f = walled_two_gels(glsm, storage=storage)
    ogs, oc, of = walled_two_gels(oracle, prec=64)
    try:
        for it in range(40):
            glsm.system_step(gs, c, f, it)
            oracle.system_step(ogs, oc, of, it)
        got = f.conservation()
        want, l1 = host_moments(of.read("f"))
        say(f"{storage}: mass excess gpu {got[0]:.3e} oracle {want[0]:.3e} (N={f.N}); momentum gpu {got[1:]} "
            f"oracle {want[1:]} L1 {l1:.3e}; u rel-L2 {rel_l2(f.read('u'), of.read('u')):.3e}")
        assert abs(got[0]) <= mass_tol * f.N
        assert np.abs(got[1:] - want[1:]).max() <= tol * l1
    finally:
        close_all(*gs, f, c, *ogs, of, oc)
