"""Writes tests/golden/cube41_one_gel.npz: the FP64 oracle's state for the small
plain-method workload -- a 41^3 D3Q19 lattice (physical box 20^3, h = 0.5),
periodic on all axes, one 6x6x6-element gel centred at (10,10,10), chemistry
at the steady state plus a splitmix64(seed 0) perturbation of amplitude 0.01,
reference constants, 25 LBM sub-steps per gel step -- after 100 and 1000
gel steps.

Stored: the gel's node positions and element chemistry (float64) at both
checkpoints, and the velocity and concentration fields at step 1000 (rounded
to float32, far below the tolerances they are compared with).  Run from the
repository root:
    python tests/golden/make_cube41_one_gel.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cube41_one_gel.npz")
CHECKPOINTS = (100, 1000)


def main():
    O.set_threads(os.cpu_count() or 1)
    g = O.Gel((6, 6, 6), (10, 10, 10))
    cp = O.Coupler([g], 64)
    fl = O.Fluid((20, 20, 20), cp)
    fl.set_periodic(1, 1, 1)
    g.perturb_chemistry(0, 0.01)
    data = {}
    t0 = time.time()
    for it in range(max(CHECKPOINTS)):
        O.system_step([g], cp, fl, it)
        n = it + 1
        if n in CHECKPOINTS:
            if n == max(CHECKPOINTS):
                data[f"u_{n}"] = fl.read("u").astype(np.float32)
                data[f"c_{n}"] = fl.read("c").astype(np.float32)
            for name in ("rn", "um", "vm", "wm", "un_norm"):
                data[f"{name}_{n}"] = g.read(name)
            print(f"step {n}: {time.time() - t0:.1f}s", flush=True)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
