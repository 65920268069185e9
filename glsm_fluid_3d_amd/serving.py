"""Serving: a request loop with continuous batching over one device.

Jobs (lattice, gels, rate constants, horizon) arrive over time.  The server
  * queues them in arrival order;
  * admits the head of the queue once its device footprint fits the HBM
    budget and the active lattice cells stay under `max_active_cells` (the
    size at which the device is saturated: more only stretches every job);
  * steps every admitted job together through one Batch (one launch per
    phase, HIP-graph captured; jobs join and leave between steps, each
    giving the same bits as when run alone);
  * streams each job's .dat files (the reference's layout and cadence)
    through an asynchronous OutputStream into its own directory;
  * retires a job the step its horizon is reached and records its arrival,
    admission and completion times;
  * checks the active jobs' state every `health_every` steps and at
    completion (non-finite populations, velocity, concentration or gel
    fields; an FP16 deviation past the storage's range is stored as Inf and
    counts) and stops a failed job at once, listing it as failed.
The loop runs on the host between chunks of steps; a chunk never passes the
nearest horizon, so completions are exact to the step.  While a chunk runs
on the device the host builds the objects of the queued jobs that already
fit (their allocations, setup kernels and uploads queue behind the chunk),
so admission after the chunk only joins them to the batch; which jobs are
admitted, and when, is the same as building them at admission.
"""
from __future__ import annotations

import os
import shutil
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import workloads as W

NSUB = 25  # LBM sub-steps per time step (fluid.cu:194-197 at the reference constants)
STORAGE_BYTES = {"f32full": 4, "f32": 4, "f16": 2, "bf16": 2}
_ALLOC_SLACK = 64 << 10  # per device allocation: allocator granularity


def gel_bytes(n: int) -> int:
    """Device bytes of one n^3-element gel (Gel's allocations, gel.cpp)."""
    L = n + 1
    nE, nN = (L + 1) ** 3, (L + 2) ** 3
    M = L ** 3 - (L - 2) ** 3
    b = 10 * nE * 8 + nN * 8          # um, um_norm, vm_norm, wm, wmp, T0, T1, T2, vol, pm; vm (node-sized)
    b += 3 * nN * 8                   # un_norm, un_robin, vn_norm
    b += 6 * nN * 24                  # rn, Fn, Fn_robin, Fdrag_robin, Veln, Vels
    b += (1 + 8 + 6) * nE * 24        # rm, rm_loc, nmSm
    b += 4 * (nN + nE + M)            # maps, boundary index
    b += 2 * 1000 * 8 + 3 * 1000 * 24 + 1000 * 300 * 24  # centre ring, filament
    return b + 40 * _ALLOC_SLACK


def fluid_bytes(edge: int, markers: int, storage: str) -> int:
    """Device bytes of a Fluid of edge^3 cells with its Coupler (fluid.cpp,
    coupling.cpp)."""
    N = edge ** 3
    plane = (N + 63) // 64 * 64
    nblk = (N + 255) // 256
    M = markers
    b = 2 * 19 * plane * STORAGE_BYTES[storage]   # ping-pong distributions
    b += N * (12 + 4 + 4 + 4 + 4 + 8)             # u, c1, c2, counts, fill cursors, ranges
    b += nblk * (4 + 4 + 32) + 64                 # flags, block sums, FP64 partials
    b += 64 * M * (8 + 12) + 4 * M                # entries, stencil tables, weight sums
    b += M * (5 * 12 + 3 * 4 + 2 * 4)             # coupler marker arrays
    return b + 30 * _ALLOC_SLACK


def output_bytes(edge: int, gel_sizes) -> int:
    """Staging buffers of the asynchronous output (device + pinned host)."""
    b = 16 * edge ** 3
    for n in gel_sizes:
        L = n + 1
        b += (9 * (L + 2) ** 3 + 6 * (L + 1) ** 3 + 11000) * 8
    return b


def job_footprint(job: W.Instance, storage: str, outputs: bool) -> int:
    sizes = [g[0] for g in job.gels]
    M = sum((n + 1) ** 3 - (n - 1) ** 3 for n in sizes)
    b = fluid_bytes(job.edge, M, storage) + sum(gel_bytes(n) for n in sizes)
    if outputs:
        b += output_bytes(job.edge, sizes)
    return b


@dataclass
class JobRecord:
    job: W.Job
    footprint: int
    admitted: float = None
    completed: float = None
    member: int = None
    objects: tuple = None
    out_dir: str = ""
    steps: int = 0
    extra: dict = field(default_factory=dict)


class Server:
    """Continuous-batching request loop (see the module docstring).

    mod: the library module (glsm_fluid_3d_amd).  hbm_budget: bytes the jobs
    may occupy (default 80% of the free device memory at start).
    out_root: directory for the jobs' .dat output (None = no output).
    keep_output: keep each finished job's directory (else removed once its
    files are written).  on_complete(record, gels, coupler, fluid) is called
    before a finished job's objects are released."""

    def __init__(self, mod, storage="f16", hbm_budget=None, max_active_cells=16_000_000, max_active_jobs=256,
                 out_root=None, keep_output=True, writers=4, max_chunk=50, autotune=True, on_complete=None,
                 health_every=500):
        self.G = mod
        self.storage = storage
        free, total = mod.device_memory()
        self.hbm_budget = int(hbm_budget if hbm_budget is not None else 0.8 * free)
        self.max_active_cells = max_active_cells
        self.max_active_jobs = max_active_jobs
        self.out_root = out_root
        self.keep_output = keep_output
        self.max_chunk = max_chunk
        self.autotune = autotune
        self.on_complete = on_complete
        self.health_every = health_every
        self.batch = mod.Batch(storage)
        self.output = mod.OutputStream(writers) if out_root else None
        if self.output:
            self.batch.set_output(self.output)
        self.queue = deque()
        self.ready = deque()  # (record, objects): the queue's head, built during the last chunk
        self.active = {}
        self.done = []
        self.rejected = []
        self.failed = []
        self.used_bytes = 0
        self.active_cells = 0
        self.cell_updates = 0
        self.steps = 0
        self.tuned = False
        # wall seconds spent admitting (joining jobs to the batch, building
        # those not built yet), stepping (until the device is idle; includes
        # the jobs built meanwhile, also counted apart) and retiring /
        # checking jobs
        self.host_s = {"admit": 0.0, "step": 0.0, "build_during_step": 0.0, "retire_and_checks": 0.0}

    # ---- admission ---------------------------------------------------------
    def submit(self, job: W.Job):
        """Queue a job; one whose footprint exceeds the whole HBM budget could
        never be admitted and would block every job behind it, so it is
        rejected at once (listed in stats()["rejected"])."""
        rec = JobRecord(job, job_footprint(job, self.storage, self.out_root is not None))
        if rec.footprint > self.hbm_budget:
            rec.extra["reason"] = f"footprint {rec.footprint} B exceeds the HBM budget {self.hbm_budget} B"
            self.rejected.append(rec)
            return False
        self.queue.append(rec)
        return True

    def _fits(self, rec: JobRecord, jobs=0, nbytes=0, cells=0) -> bool:
        """Whether `rec` fits beside the active jobs changed by `jobs` jobs
        holding `nbytes` and `cells` (negative: jobs about to leave)."""
        n_active = len(self.active) + jobs
        if n_active >= self.max_active_jobs:
            return False
        if self.used_bytes + nbytes + rec.footprint > self.hbm_budget:
            return False
        return n_active == 0 or self.active_cells + cells + rec.job.cells <= self.max_active_cells

    def _prebuild(self, leaving=()):
        """Build the objects of the queued jobs that will fit once the jobs
        in `leaving` (those reaching their horizon with the running chunk)
        are retired, beside those already built, in queue order (called while
        the chunk runs).  The leaving jobs' device memory still counts: it is
        freed only after the chunk, and the new objects allocate now."""
        jobs = len(self.ready) - len(leaving)
        nbytes = sum(r.footprint for r, _ in self.ready)
        cells = sum(r.job.cells for r, _ in self.ready) - sum(r.job.cells for r in leaving)
        while self.queue and self._fits(self.queue[0], jobs, nbytes, cells):
            rec = self.queue.popleft()
            self.ready.append((rec, W.build(self.G, rec.job, self.storage)))
            jobs += 1
            nbytes += rec.footprint
            cells += rec.job.cells

    def _next_admissible(self):
        """The queue's head (built or not) if it fits, else None."""
        if self.ready:
            return self.ready[0] if self._fits(self.ready[0][0]) else None
        if self.queue and self._fits(self.queue[0]):
            return self.queue[0], None
        return None

    def _admit(self, now: float) -> int:
        n = 0
        while True:
            head = self._next_admissible()
            if head is None:
                break
            rec, objs = head
            if objs is None:
                self.queue.popleft()
                objs = W.build(self.G, rec.job, self.storage)
            else:
                self.ready.popleft()
            if self.out_root:
                rec.out_dir = os.path.join(self.out_root, f"job{rec.job.job_id}")
                os.makedirs(rec.out_dir, exist_ok=True)
            rec.member = self.batch.add(*objs, iteration=0, write_files=self.out_root is not None,
                                        out_dir=rec.out_dir)
            rec.objects = objs
            rec.admitted = now
            self.active[rec.member] = rec
            self.used_bytes += rec.footprint
            self.active_cells += rec.job.cells
            n += 1
        if n and self.autotune and not self.tuned and self.active_cells >= self.max_active_cells // 2:
            self.batch.autotune(reps=2)
            self.tuned = True
        return n

    # ---- completion --------------------------------------------------------
    def _health(self, rec):
        """None if the job's state is finite, else a description of what is not."""
        gels, c, f = rec.objects
        h = f.check_finite()
        bad_gel = sum(g.check_finite() for g in gels)
        if h["limit"] > 0:
            rec.extra["range_used"] = max(rec.extra.get("range_used", 0.0), h["max_abs_dev"] / h["limit"])
        if h["bad_cells"] or bad_gel:
            return (f"non-finite state at step {self.batch.iteration(rec.member)}: {h['bad_cells']} lattice cells, "
                    f"{bad_gel} gel values")
        return None

    def _release(self, rec, now, failure=None):
        self.batch.remove(rec.member)
        gels, c, f = rec.objects
        if self.output:
            for o in gels + [f]:
                self.output.release(o)
        if self.on_complete and failure is None:
            self.on_complete(rec, gels, c, f)
        for o in gels + [f, c]:
            o.close()
        rec.objects = None
        rec.completed = now
        del self.active[rec.member]
        self.used_bytes -= rec.footprint
        self.active_cells -= rec.job.cells
        if failure is None:
            self.done.append(rec)
        else:
            rec.extra["failure"] = failure
            self.failed.append(rec)
        if self.out_root and not self.keep_output:
            shutil.rmtree(rec.out_dir, ignore_errors=True)

    def _retire(self, now: float, check_all: bool = False):
        for rec in list(self.active.values()):
            it = self.batch.iteration(rec.member)
            finished = it >= rec.job.horizon
            if not (finished or check_all):
                continue
            failure = self._health(rec)
            if failure is not None or finished:
                rec.steps = it
                self._release(rec, now, failure)

    # ---- the loop ----------------------------------------------------------
    def run(self, jobs, duration=None, clock=None, idle_sleep=0.002, progress_every=None, log=None,
            until_completed=None):
        """Serve `jobs` (arrival times relative to the start) until all are
        done, `duration` seconds have passed, or (with until_completed) that
        many jobs have completed.  `clock` () -> seconds since start
        (default: wall time).  With progress_every, `log` (default: stderr)
        gets a status line every that many seconds."""
        import sys

        log = log or (lambda m: print(m, file=sys.stderr, flush=True))
        next_report = progress_every
        t0 = time.perf_counter()
        now_fn = clock or (lambda: time.perf_counter() - t0)
        pending = deque(sorted(jobs, key=lambda j: j.arrival))
        self.t_start = now_fn()
        while True:
            now = now_fn()
            if duration is not None and now >= duration:
                break
            if until_completed is not None and len(self.done) >= until_completed:
                break
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            t_a = time.perf_counter()
            self._admit(now)
            self.host_s["admit"] += time.perf_counter() - t_a
            if not self.active:
                if not pending and not self.queue and not self.ready:
                    break
                if clock is None:
                    time.sleep(idle_sleep)
                else:
                    clock.advance_to(pending[0].arrival if pending else now)
                continue
            left = min(r.job.horizon - self.batch.iteration(r.member) for r in self.active.values())
            n = max(1, min(self.max_chunk, left))
            t_s = time.perf_counter()
            self.batch.step(n)
            t_b = time.perf_counter()
            # (iteration counters advance when the steps are enqueued)
            self._prebuild([r for r in self.active.values() if self.batch.iteration(r.member) >= r.job.horizon])
            self.host_s["build_during_step"] += time.perf_counter() - t_b
            self.G.synchronize()
            self.host_s["step"] += time.perf_counter() - t_s
            self.steps += n
            self.cell_updates += self.active_cells * NSUB * n
            if clock is not None:
                clock.advance(n)
            check = self.health_every and (self.steps // self.health_every) != ((self.steps - n) // self.health_every)
            t_r = time.perf_counter()
            self._retire(now_fn(), check_all=bool(check))
            self.host_s["retire_and_checks"] += time.perf_counter() - t_r
            if progress_every and now >= next_report:
                next_report += progress_every
                log(f"[serve {now:7.1f} s] active {len(self.active)} queued {len(self.queue)} "
                    f"done {len(self.done)} cells {self.active_cells} "
                    f"MLUPS {self.cell_updates / max(now - self.t_start, 1e-9) / 1e6:.0f}")
        self.elapsed = now_fn() - self.t_start
        if self.output:
            self.output.flush()
        return self.stats()

    def stats(self):
        lat = np.array([r.completed - r.job.arrival for r in self.done]) if self.done else np.zeros(0)
        wait = np.array([r.admitted - r.job.arrival for r in self.done]) if self.done else np.zeros(0)
        pct = lambda a, q: float(np.percentile(a, q)) if a.size else None
        out = {
            "completed": len(self.done), "active": len(self.active), "queued": len(self.queue) + len(self.ready),
            "rejected": [r.job.job_id for r in self.rejected],
            "failed": {r.job.job_id: r.extra["failure"] for r in self.failed},
            "latency_p50_s": pct(lat, 50), "latency_p99_s": pct(lat, 99),
            "queue_wait_p50_s": pct(wait, 50), "queue_wait_p99_s": pct(wait, 99),
            "mlups": self.cell_updates / max(self.elapsed, 1e-9) / 1e6 if hasattr(self, "elapsed") else None,
            "elapsed_s": getattr(self, "elapsed", None), "batch_steps": self.steps,
            "hbm_budget_GB": self.hbm_budget / 1e9,
            "wall_s_by_phase": {k: round(v, 3) for k, v in self.host_s.items()},
        }
        if self.output:
            out["output"] = self.output.status()
        return out

    def close(self):
        for rec in list(self.active.values()):
            self.batch.remove(rec.member)
            gels, c, f = rec.objects
            if self.output:
                for o in gels + [f]:
                    self.output.release(o)
            for o in gels + [f, c]:
                o.close()
        self.active.clear()
        while self.ready:
            gels, c, f = self.ready.popleft()[1]
            for o in gels + [f, c]:
                o.close()
        if self.output:
            self.output.close()
        self.batch.close()


class StepClock:
    """A virtual clock for tests: time advances by `dt` per batch step, or
    jumps to the next arrival when the server is idle."""

    def __init__(self, dt=1.0):
        self.t = 0.0
        self.dt = dt

    def __call__(self):
        return self.t

    def advance(self, steps):
        self.t += steps * self.dt

    def advance_to(self, t):
        self.t = max(self.t, t)
