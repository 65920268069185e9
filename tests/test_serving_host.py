"""The serving loop's host logic on the CPU, with a stand-in for the device
library (no GPU): admission in queue order under the HBM budget and the
active-cell cap, completion exactly at each job's horizon, rejection of a job
that can never fit, and -- with objects built while a chunk runs -- the same
admissions, at the same times, as building them at admission."""
import numpy as np
import pytest

from glsm_fluid_3d_amd import serving as S
from glsm_fluid_3d_amd import workloads as W


class _Obj:
    """Gel / coupler / fluid stand-in: records construction and closing."""

    made = []

    def __init__(self, kind, *a, **k):
        self.kind, self.closed = kind, False
        _Obj.made.append(self)

    def set_chemistry_rates(self, *a):
        pass

    def perturb_chemistry(self, *a):
        pass

    def set_storage(self, *a):
        pass

    def check_finite(self):
        return {"bad_cells": 0, "max_abs_dev": 0.0, "limit": 0.0} if self.kind == "fluid" else 0

    def close(self):
        assert not self.closed, "closed twice"
        self.closed = True


class _Batch:
    def __init__(self, storage):
        self.members, self.next_id, self.steps_run = {}, 1, 0

    def add(self, gels, c, f, iteration=0, write_files=False, out_dir=""):
        assert not any(o.closed for o in gels + [c, f])
        self.members[self.next_id] = [iteration, (gels, c, f)]
        self.next_id += 1
        return self.next_id - 1

    def remove(self, mid):
        del self.members[mid]

    def iteration(self, mid):
        return self.members[mid][0]

    def step(self, n):
        assert self.members
        for m in self.members.values():
            m[0] += n
        self.steps_run += n

    def autotune(self, reps=2):
        return {}, 0.0

    def close(self):
        pass


class _Lib:
    """The parts of glsm_fluid_3d_amd the server uses."""

    def __init__(self, free_bytes):
        self.free = free_bytes

    def device_memory(self):
        return self.free, self.free

    def Batch(self, storage):
        return _Batch(storage)

    def Gel(self, *a, **k):
        return _Obj("gel")

    def Coupler(self, gels):
        return _Obj("coupler")

    def Fluid(self, box, c):
        return _Obj("fluid")

    def synchronize(self):
        pass


def _jobs(n=40):
    return W.poisson_jobs(seed=7, rate=20.0, duration=60.0, edge_range=(41, 161), horizon_range=(40, 400),
                          max_jobs=n)


def _serve(lib, jobs, prebuild=True, **kw):
    srv = S.Server(lib, storage="f16", autotune=False, **kw)
    srv.built_ahead = 0
    if not prebuild:
        srv._prebuild = lambda leaving=(): None
    else:
        inner = srv._prebuild

        def counted(leaving=()):
            n = len(srv.ready)
            inner(leaving)
            srv.built_ahead += len(srv.ready) - n

        srv._prebuild = counted
    clock = S.StepClock(dt=0.05)
    st = srv.run(jobs, clock=clock)
    rec = sorted((r.job.job_id, r.admitted, r.completed, r.steps) for r in srv.done)
    srv.close()
    return st, rec, srv


@pytest.mark.parametrize("cap", [4_000_000, 16_000_000])
def test_build_ahead_admits_the_same_jobs_at_the_same_times(cap):
    jobs = _jobs()
    budget = 40 * S.job_footprint(jobs[0], "f16", False)
    st0, rec0, _ = _serve(_Lib(budget / 0.8), jobs, prebuild=False, max_active_cells=cap)
    st1, rec1, srv = _serve(_Lib(budget / 0.8), jobs, prebuild=True, max_active_cells=cap)
    assert st0["completed"] == st1["completed"] == len(jobs)
    assert rec0 == rec1
    # every job ran exactly to its horizon, and something was built ahead
    assert all(steps == next(j.horizon for j in jobs if j.job_id == jid) for jid, _, _, steps in rec1)
    assert srv.built_ahead > 0


def test_admission_respects_the_cell_cap_and_queue_order():
    jobs = _jobs(30)
    cap = 3_000_000
    lib = _Lib(1e15)
    srv = S.Server(lib, storage="f16", autotune=False, max_active_cells=cap)
    clock = S.StepClock(dt=0.05)
    seen_over = []
    orig_step = srv.batch.step

    def step(n):
        cells = sum(r.job.cells for r in srv.active.values())
        if len(srv.active) > 1 and cells > cap:
            seen_over.append(cells)
        orig_step(n)

    srv.batch.step = step
    srv.run(jobs, clock=clock)
    assert not seen_over
    admitted = [r.job.job_id for r in sorted(srv.done, key=lambda r: (r.admitted, r.job.job_id))]
    assert admitted == sorted(admitted), "jobs are admitted in arrival order"
    srv.close()


def test_oversize_job_rejected_and_objects_closed():
    jobs = _jobs(12)
    big = W.Job(edge=401, gels=[], seed=1, job_id=999, horizon=10, arrival=0.01)
    budget = 2 * max(S.job_footprint(j, "f16", False) for j in jobs)
    _Obj.made.clear()
    st, rec, srv = _serve(_Lib(budget / 0.8), jobs + [big])
    assert st["rejected"] == [999]
    assert st["completed"] == len(jobs)
    # each admitted job's objects were built once and closed once
    assert len(_Obj.made) == sum(len(j.gels) + 2 for j in jobs)
    assert all(o.closed for o in _Obj.made)


def test_step_clock_and_horizons_exact():
    jobs = _jobs(10)
    st, rec, _ = _serve(_Lib(1e15), jobs)
    horizons = {j.job_id: j.horizon for j in jobs}
    for jid, admitted, completed, steps in rec:
        assert steps == horizons[jid]
        # the virtual clock advances 0.05 s per step: a job's service time
        # is at least its horizon in steps
        assert completed - admitted >= 0.05 * steps - 1e-9
    assert np.isfinite([r[2] for r in rec]).all()


def test_close_releases_active_and_built_ahead_jobs():
    jobs = _jobs(40)
    budget = 40 * S.job_footprint(jobs[0], "f16", False)
    _Obj.made.clear()
    srv = S.Server(_Lib(budget / 0.8), storage="f16", autotune=False, max_active_cells=4_000_000)
    srv.run(jobs, duration=3.0, clock=S.StepClock(dt=0.05))
    assert srv.active and srv.queue
    srv._prebuild(list(srv.active.values()))  # as if a chunk retiring every active job were running
    assert srv.ready
    srv.close()
    assert _Obj.made and all(o.closed for o in _Obj.made)
