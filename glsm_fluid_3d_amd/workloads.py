"""Seeded synthetic workloads: instance and job descriptions (no device work).

Lattice sizes follow the reference's grid rule (fluid.cu:167-197): a
physical box of side b gives int(b / h) + 1 nodes with h = 0.5, so an odd
lattice edge E is the box (E - 1) / 2.  Gels are cubes of n elements (dx =
1, gel.cu:633-652) centred at the given position.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

F_REF, EP_REF = 0.9, 0.3  # gel chemistry rate constants (gel.cu:633-652)


def round_to_odd(x: float) -> int:
    return int(2 * round((x - 1) / 2) + 1)


def place_gels(rng, box: float, sizes, margin: float = 2.0, tries: int = 200, room: float = 0.0):
    """Seeded centres for cube gels of the given element counts inside a
    cubic box: at least `margin` + room * n from the walls and
    1 + 2 * room * n apart from each other (their boxes separated along some
    axis).  `room` is the fraction of its edge a gel may grow by on each
    side.  A gel that finds no place in `tries` draws is dropped.  Returns
    [(centre, n)]."""
    out = []
    for n in sizes:
        half = n / 2.0
        lo, hi = half + margin + room * n, box - half - margin - room * n
        if hi <= lo:
            continue
        for _ in range(tries):
            p = rng.uniform(lo, hi, 3)
            if all(np.any(np.abs(p - q) >= half + m / 2.0 + 1.0 + room * (n + m)) for q, m in out):
                out.append((p, n))
                break
    return out


@dataclass
class Instance:
    """One gel-fluid instance: cubic lattice of `edge` nodes, gels as
    (n elements, centre, f, ep), perturbation seed."""
    edge: int
    gels: list
    seed: int

    @property
    def box(self) -> float:
        return (self.edge - 1) * 0.5

    @property
    def cells(self) -> int:
        return self.edge ** 3


def batch_instances(count=64, seed=3, edge_range=(33, 97), gels_range=(1, 4), n_range=(4, 8), rate_jitter=0.2):
    """The batching workload: lattice edge log-uniform in edge_range rounded
    to odd, 1-4 cube gels of 4^3-8^3 elements at seeded non-overlapping
    positions, rate constants f, ep uniform within +-20% of the reference's."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        edge = round_to_odd(np.exp(rng.uniform(np.log(edge_range[0]), np.log(edge_range[1]))))
        box = (edge - 1) * 0.5
        ng = int(rng.integers(gels_range[0], gels_range[1] + 1))
        sizes = [int(rng.integers(n_range[0], n_range[1] + 1)) for _ in range(ng)]
        placed = place_gels(rng, box, sizes)
        gels = [(n, tuple(float(v) for v in p), F_REF * rng.uniform(1 - rate_jitter, 1 + rate_jitter),
                 EP_REF * rng.uniform(1 - rate_jitter, 1 + rate_jitter)) for p, n in placed]
        out.append(Instance(edge, gels, seed * 1000 + k))
    return out


@dataclass
class Job(Instance):
    """A simulation request: an instance, a horizon in time steps and its
    arrival time (seconds from the start of the stream)."""
    job_id: int = 0
    horizon: int = 0
    arrival: float = 0.0
    extra: dict = field(default_factory=dict)


# Over the serving horizons (up to 20000 steps) the BZ chemistry swells a gel
# of the reference's constants until its element volume is ~6x the initial
# one: each face moves out by up to about half the gel's edge.  A gel whose
# surface reaches a wall puts IBM markers where their stencils are clipped
# to a few cells; the spread force then concentrates (it is normalised by
# the clipped weight sum, fluid_kernels.cu:300-348) and the explicit
# coupling diverges once the ramped stiffness is large enough -- the same
# happens in FP32 and FP64.  Serving jobs therefore keep that much room
# around each gel.
SWELL_ROOM = 0.5


def poisson_jobs(seed=4, rate=20.0, duration=600.0, edge_range=(41, 161), gels_range=(1, 6), n=6,
                 horizon_range=(1000, 20000), max_jobs=None, room=SWELL_ROOM):
    """The serving workload: Poisson arrivals (mean `rate` jobs/s) over
    `duration` seconds; lattice edge log-uniform in edge_range rounded to
    odd, 1-6 gels of n^3 elements at seeded non-overlapping positions with
    room to swell (place_gels, `room`), the reference's rate constants,
    horizon uniform in horizon_range.  A gel that finds no place in a small
    box is dropped, so small lattices hold fewer gels."""
    rng = np.random.default_rng(seed)
    jobs = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration or (max_jobs is not None and len(jobs) >= max_jobs):
            break
        edge = round_to_odd(np.exp(rng.uniform(np.log(edge_range[0]), np.log(edge_range[1]))))
        box = (edge - 1) * 0.5
        ng = int(rng.integers(gels_range[0], gels_range[1] + 1))
        placed = place_gels(rng, box, [n] * ng, room=room)
        gels = [(m, tuple(float(v) for v in p), F_REF, EP_REF) for p, m in placed]
        horizon = int(rng.integers(horizon_range[0], horizon_range[1] + 1))
        k = len(jobs)
        jobs.append(Job(edge=edge, gels=gels, seed=seed * 100000 + k, job_id=k + 1, horizon=horizon, arrival=t))
    return jobs


def build(mod, inst: Instance, storage=None):
    """The instance's objects through `mod` (the library or the oracle):
    gels (rates set, chemistry perturbed by 0.01 from the seed), coupler,
    fluid (walls)."""
    gels = []
    for j, (n, pos, f, ep) in enumerate(inst.gels):
        g = mod.Gel((n, n, n), pos, gel_id=j + 1)
        if (f, ep) != (F_REF, EP_REF):
            g.set_chemistry_rates(f, ep)
        g.perturb_chemistry(inst.seed * 8 + j, 0.01)
        gels.append(g)
    c = mod.Coupler(gels)
    b = int(round(inst.box))
    f = mod.Fluid((b, b, b), c)
    if storage:
        f.set_storage(storage)
    return gels, c, f


# ---- the fixed single-lattice workloads -----------------------------------
def two_gels_sim_cpp(mod, storage=None, prec=None, seed=1, box=50, n=6):
    """The reference default scene: a cubic box of side `box` (101^3 nodes for
    50) with walls, two n^3-element gels at x = 1/3 and 2/3 of the box,
    centred in y and z (sim.cpp's fractional placement, sim.cpp:26-30),
    chemistry perturbed by 0.01 from splitmix64(seed + gel index).  `prec`
    selects the oracle's fluid arithmetic (ignored by the library)."""
    gels = [mod.Gel((n, n, n), (box / 3.0, box / 2.0, box / 2.0), gel_id=1),
            mod.Gel((n, n, n), (2.0 * box / 3.0, box / 2.0, box / 2.0), gel_id=2)]
    for i, g in enumerate(gels):
        g.perturb_chemistry(seed + i, 0.01)
    c = mod.Coupler(gels) if prec is None else mod.Coupler(gels, prec)
    f = mod.Fluid((box, box, box), c)
    if storage:
        f.set_storage(storage)
    return gels, c, f


def four_gel_centres(seed=2, box=100.0, n=8, count=4):
    """Seeded uniform centres of n^3-element gels in a cubic box, at least 3
    from the walls, rejected when their boxes come closer than one element
    (the 201^3 precision workload)."""
    rng = np.random.default_rng(seed)
    half = n / 2.0
    out = []
    while len(out) < count:
        p = rng.uniform(half + 3, box - half - 3, 3)
        if all(np.any(np.abs(p - q) >= n + 1) for q in out):
            out.append(p)
    return [tuple(float(v) for v in p) for p in out]


def four_gels_201(mod, storage=None, prec=None):
    """201^3 nodes (box 100) with walls, four 8^3-element gels at
    four_gel_centres(), chemistry perturbed by 0.01 from splitmix64(2 + gel
    index)."""
    gels = [mod.Gel((8, 8, 8), p, gel_id=i + 1) for i, p in enumerate(four_gel_centres())]
    for i, g in enumerate(gels):
        g.perturb_chemistry(2 + i, 0.01)
    c = mod.Coupler(gels) if prec is None else mod.Coupler(gels, prec)
    f = mod.Fluid((100, 100, 100), c)
    if storage:
        f.set_storage(storage)
    return gels, c, f
