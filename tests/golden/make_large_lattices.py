"""Writes the FP64 oracle's fingerprints of the two large single-lattice
workloads, for the GPU parity tests of the plain and precision forms:

  walled101_two_gels_ref.npz   101^3 nodes (box 50), walls, two 6^3-element
                               gels at x = 1/3, 2/3 of the box (sim.cpp's
                               placement), chemistry perturbed from seeds 1, 2;
                               the reference's IBM ramp (20000 steps);
                               checkpoints after 100 and 200 gel steps
  walled101_two_gels_full.npz  the same scene with the IBM ramp cut to 20
                               steps (IBM_RAMP_STEPS=20: full stiffness from
                               step 20 on); 100 and 200 steps
  walled201_four_gels_full.npz 201^3 nodes (box 100), walls, four 8^3-element
                               gels at seeded non-overlapping centres, ramp 20;
                               50 and 100 steps
  walled201_four_gels_ref.npz  the same scene with the reference's IBM ramp;
                               50 and 100 steps

Each checkpoint holds, from the oracle's FP64 state:
  * u, c and rho - 1 (rho = sum of the next sub-step's pre-collision
    populations) on a fixed cell sample: for every gel, the z-plane through
    its centre within +-W nodes of it in x and y (W = 50 for 101^3, so the
    whole plane; W = 24 for 201^3), plus 16384 seeded random cells; stored as
    float32 (its rounding, ~6e-8 relative, is far below every tolerance these
    are compared with);
  * full-field FP64 scalars: sum(rho - 1), sum(rho u) = sum_q f_q c_q, |u|_2,
    |c|_2, sum(c);
  * every gel's node positions rn and element chemistry um, vm, wm (FP64).

Full fields would be 12-100 MB per checkpoint; the sample keeps each fixture
near 1 MB while still giving a relative L2 over ~26k cells, a third of them
in the planes where the gels force the fluid.

Run from the repository root (each 201^3 case takes about half an hour on 8 cores):
    python tests/golden/make_large_lattices.py walled101_two_gels_ref
    python tests/golden/make_large_lattices.py walled101_two_gels_full
    python tests/golden/make_large_lattices.py walled201_four_gels_full
    python tests/golden/make_large_lattices.py walled201_four_gels_ref
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402
from glsm_fluid_3d_amd import workloads as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0], [1, 0, 1], [-1, 0, 1], [-1, 0, -1],
              [1, 0, -1], [0, 1, 1], [0, 1, -1], [0, -1, -1], [0, -1, 1]], dtype=np.float64)
SAMPLE_SEED, SAMPLE_CELLS = 12345, 16384

CASES = {
    "walled101_two_gels_ref": dict(scene="two", ramp=None, checkpoints=(100, 200), window=50),
    "walled101_two_gels_full": dict(scene="two", ramp=20.0, checkpoints=(100, 200), window=50),
    "walled201_four_gels_full": dict(scene="four", ramp=20.0, checkpoints=(50, 100), window=24),
    "walled201_four_gels_ref": dict(scene="four", ramp=None, checkpoints=(50, 100), window=24),
}


def scene(mod, name, prec=None, storage=None):
    """The objects of a case through `mod` (the oracle or the library)."""
    cfg = CASES[name]
    if cfg["scene"] == "two":
        gels, c, f = W.two_gels_sim_cpp(mod, storage=storage, prec=prec)
    else:
        gels, c, f = W.four_gels_201(mod, storage=storage, prec=prec)
    if cfg["ramp"] is not None:
        f.set_ibm(1.0, cfg["ramp"])
    return gels, c, f


def gel_centres(name):
    if CASES[name]["scene"] == "four":
        return W.four_gel_centres()
    return [(50 / 3.0, 25.0, 25.0), (100 / 3.0, 25.0, 25.0)]


def sample_index(name, grid):
    """Flat indices of the compared cells (sorted, unique): the windowed
    centre plane of every gel and seeded random cells."""
    Nx, Ny, Nz = grid
    W_ = CASES[name]["window"]
    idx = []
    for p in gel_centres(name):
        x0, y0, z0 = (int(round(v / 0.5)) for v in p)
        xs = np.arange(max(0, x0 - W_), min(Nx, x0 + W_ + 1))
        ys = np.arange(max(0, y0 - W_), min(Ny, y0 + W_ + 1))
        yy, xx = np.meshgrid(ys, xs, indexing="ij")
        idx.append((xx + Nx * yy + Nx * Ny * z0).ravel())
    rng = np.random.default_rng(SAMPLE_SEED)
    idx.append(rng.choice(Nx * Ny * Nz, SAMPLE_CELLS, replace=False))
    return np.unique(np.concatenate(idx)).astype(np.int64)


def fingerprint(gels, cells, rho_dev, u, c):
    """Dict of the checkpoint's sampled fields, full-field norms and gels."""
    out = {"u": u[cells].astype(np.float32), "c": c[cells].astype(np.float32),
           "rho": rho_dev[cells].astype(np.float32),
           "norm_u": float(np.linalg.norm(u)), "norm_c": float(np.linalg.norm(c)), "sum_c": float(c.sum())}
    for k, g in enumerate(gels):
        for name in ("rn", "um", "vm", "wm"):
            out[f"gel{k}_{name}"] = g.read(name)
    return out


def oracle_moments(f):
    """(rho - 1 per cell, [sum(rho - 1), sum(rho u)]) of the oracle's
    pre-collision populations, in FP64."""
    fp = f.read("f")  # [19, N]
    rho_dev = fp.sum(axis=0) - 1.0
    mom = C.T @ fp  # [3, N]
    return rho_dev, np.array([rho_dev.sum(), mom[0].sum(), mom[1].sum(), mom[2].sum()])


def main(name):
    cfg = CASES[name]
    O.build()
    O.set_threads(os.cpu_count() or 1)
    gels, c, f = scene(O, name, prec=64)
    cells = sample_index(name, f.grid)
    data = {"cells": cells, "grid": np.array(f.grid, dtype=np.int32)}
    t0 = time.time()
    for it in range(max(cfg["checkpoints"])):
        O.system_step(gels, c, f, it)
        n = it + 1
        if n % 10 == 0:
            print(f"{name}: step {n} ({time.time() - t0:.0f} s)", flush=True)
        if n in cfg["checkpoints"]:
            rho_dev, cons = oracle_moments(f)
            fp = fingerprint(gels, cells, rho_dev, f.read("u"), f.read("c"))
            fp["conservation"] = cons
            for k, v in fp.items():
                data[f"{k}_{n}"] = v
    out = os.path.join(HERE, name + ".npz")
    np.savez_compressed(out, **data)
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    for arg in sys.argv[1:] or sorted(CASES):
        main(arg)
