"""Failure detection and the FP64 density field.

  * check_finite on a healthy lattice reports no bad cell, and its largest
    |f_q - w_q| equals the host's maximum over the populations written;
  * a NaN written into one population (through glsm_fluid_write) is found,
    and still found after it has spread for a sub-step;
  * FP16 deviation storage: a deviation beyond 65504 / 2^s is stored as Inf
    (the storage's saturation), and the scan reports it; `limit` is the
    storage's range;
  * a NaN written into a gel's um is found by the gel scan;
  * serving: a job whose state turns non-finite is stopped and listed as
    failed, and the other jobs complete with the same bits as alone;
  * the rho field (rho - 1 summed in FP64 from the stored values) equals
    the same sum taken on the host, bit for bit, for FP32 and FP32-deviation
    storage.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float32)


def lattice(G, storage="f32", box=(8, 6, 7)):
    g = G.Gel((2, 2, 2), (4.0, 3.0, 3.5), gel_id=1)
    c = G.Coupler([g])
    f = G.Fluid(box, c)
    f.set_storage(storage)
    return g, c, f


def close_all(*objs):
    for o in objs:
        o.close()


def perturbed_populations(f, amp, seed=5):
    rng = np.random.default_rng(seed)
    return (W[:, None] + amp * rng.standard_normal((19, f.N))).astype(np.float32)


@pytest.mark.parametrize("storage", ["f32", "f32full", "f16", "bf16"])
def test_check_finite_healthy_and_nan(glsm, storage):
    G = glsm
    g, c, f = lattice(G, storage)
    try:
        pops = perturbed_populations(f, 1e-3)
        f.write("f", pops)
        h = f.check_finite()
        print(storage, h, flush=True)
        assert h["bad_cells"] == 0
        host_max = np.abs(pops.astype(np.float64) - W[:, None].astype(np.float64)).max()
        rel = {"f32": 1e-6, "f32full": 1e-4, "f16": 2e-3, "bf16": 1e-2}[storage]
        assert abs(h["max_abs_dev"] - host_max) <= rel * host_max
        assert g.check_finite() == 0
        bad = pops.copy()
        bad[7, 123] = np.nan
        f.write("f", bad)
        assert f.check_finite()["bad_cells"] == 1
        G.system_step([g], c, f, 0)  # 25 sub-steps: the NaN spreads
        h = f.check_finite()
        assert h["bad_cells"] > 1
    finally:
        close_all(g, f, c)


def test_fp16_saturation_is_reported(glsm):
    G = glsm
    g, c, f = lattice(G, "f16")
    try:
        h = f.check_finite()
        assert h["limit"] == pytest.approx(65504.0 * 2.0 ** -20)
        pops = perturbed_populations(f, 1e-4)
        pops[3, 42] = W[3] + 0.07  # deviation above 65504 * 2^-20 = 0.0625
        f.write("f", pops)
        h = f.check_finite()
        assert h["bad_cells"] == 1
        pops[3, 42] = W[3] + 0.06  # in range
        f.write("f", pops)
        h = f.check_finite()
        assert h["bad_cells"] == 0 and h["max_abs_dev"] == pytest.approx(0.06, rel=1e-3)
        assert h["max_abs_dev"] / h["limit"] > 0.9
    finally:
        close_all(g, f, c)


def test_gel_nan_is_reported(glsm):
    G = glsm
    g, c, f = lattice(G)
    try:
        um = g.read("um")
        um[10] = np.nan
        g.write("um", um)
        assert g.check_finite() == 1
        rn = g.read("rn")
        rn[5, 1] = np.inf
        g.write("rn", rn)
        assert g.check_finite() == 2
    finally:
        close_all(g, f, c)


@pytest.mark.parametrize("storage", ["f32", "f32full"])
def test_rho_field_is_the_fp64_sum_of_the_stored_values(glsm, storage):
    G = glsm
    g, c, f = lattice(G, storage)
    try:
        pops = perturbed_populations(f, 1e-3, seed=9)
        f.write("f", pops)
        rho = f.read("rho")
        # the stored values: f itself, or d = f - w rounded to FP32
        stored = pops if storage == "f32full" else (pops - W[:, None]).astype(np.float32)
        acc = np.full(f.N, -1.0 if storage == "f32full" else 0.0)
        for q in range(19):  # the kernel's order: q = 0 .. 18
            acc = acc + stored[q].astype(np.float64)
        np.testing.assert_array_equal(rho, acc)
    finally:
        close_all(g, f, c)


def test_serving_stops_a_failed_job(glsm):
    G = glsm
    from glsm_fluid_3d_amd import serving as S
    from glsm_fluid_3d_amd import workloads as W_

    jobs = [W_.Job(edge=21, gels=[(4, (5.0, 5.0, 5.0), 0.9, 0.3)], seed=60 + k, job_id=k + 1, horizon=9,
                   arrival=float(k)) for k in range(3)]
    done = {}

    def keep(rec, gels, c, f):
        done[rec.job.job_id] = (f.read("fpost"), gels[0].read("rn"))

    srv = S.Server(G, storage="f32", max_active_cells=10 ** 9, max_chunk=2, autotune=False, on_complete=keep,
                   health_every=2)
    orig = srv._admit

    def admit(now):
        n = orig(now)
        for rec in srv.active.values():
            if rec.job.job_id == 2 and "poisoned" not in rec.extra:
                gels, c, f = rec.objects
                pops = f.read("f")
                pops[4, 100] = np.nan
                f.write("f", pops)
                rec.extra["poisoned"] = True
        return n

    srv._admit = admit
    try:
        stats = srv.run(jobs, clock=S.StepClock(1.0))
    finally:
        srv.close()
    print(stats, flush=True)
    assert stats["completed"] == 2 and list(stats["failed"]) == [2]
    assert "non-finite" in stats["failed"][2]
    for job in (jobs[0], jobs[2]):
        gels, c, f = W_.build(G, job, "f32")
        try:
            for it in range(job.horizon):
                G.system_step(gels, c, f, it)
            np.testing.assert_array_equal(done[job.job_id][0], f.read("fpost"))
            np.testing.assert_array_equal(done[job.job_id][1], gels[0].read("rn"))
        finally:
            close_all(*gels, f, c)


def test_serving_rejects_a_job_larger_than_the_budget(glsm):
    G = glsm
    from glsm_fluid_3d_amd import serving as S
    from glsm_fluid_3d_amd import workloads as W_

    small = [W_.Job(edge=21, gels=[(4, (5.0, 5.0, 5.0), 0.9, 0.3)], seed=70 + k, job_id=k + 2, horizon=3,
                    arrival=1.0) for k in range(2)]
    big = W_.Job(edge=101, gels=[(6, (25.0, 25.0, 25.0), 0.9, 0.3)], seed=1, job_id=1, horizon=3, arrival=0.0)
    budget = 3 * S.job_footprint(small[0], "f32", outputs=False)
    srv = S.Server(G, storage="f32", hbm_budget=budget, max_chunk=2, autotune=False)
    try:
        stats = srv.run([big] + small, clock=S.StepClock(1.0))
    finally:
        srv.close()
    assert stats["rejected"] == [1] and stats["completed"] == 2 and stats["queued"] == 0
