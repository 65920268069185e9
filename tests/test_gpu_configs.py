"""The large workloads on the GPU, against the FP64 restatement.

Single lattices (the plain and precision forms), against the oracle's
fingerprints in tests/golden/ (written by tests/golden/make_large_lattices.py):
  * walled101_two_gels_ref / _full: 101^3 nodes with walls, two 6^3 gels
    placed as in sim.cpp, the reference's IBM ramp (20000 steps) and a
    20-step ramp (full IBM stiffness from step 20), 200 steps;
  * walled201_four_gels_ref / _full: 201^3 nodes with walls, four 8^3
    gels at seeded non-overlapping centres, the reference's IBM ramp and a
    20-step ramp, 100 steps.
Per storage (FP32 deviations, FP16 and BF16 deviations with FP64 moments)
the relative L2 errors of u, c and rho - 1 over the fixture's ~26k sampled
cells, the largest node-position error and the FP64 mass / momentum sums
are printed; tolerances, at every checkpoint:
  FP32: u <= 1e-5 (2e-5 on the 201^3 lattice with the reference's ramp,
        where the velocity signal is 1e-6 at step 50), c <= 3e-6, rho - 1 <=
        1e-5, rn <= 5e-7 gel units (c reaches 1.1e-6 on the 201^3 lattice by
        step 50);
  FP16: u <= 1e-3 (the precision form's velocity target) with the reference's
        IBM ramp; 2e-3 at full IBM stiffness from step 20 (measured 1.1e-3 at
        101^3, 1.5e-3 at 201^3), c <= 1e-4, rho - 1 <= 2e-3;
  BF16: u <= 1e-2 (2e-2 at full stiffness: 1.1e-2 at 101^3, 1.5e-2 at
        201^3), c <= 1e-3, rho - 1 <= 2e-2 (bfloat16's 8-bit significand
        misses the 1e-3 velocity bound);
  mass sum(rho - 1) per cell within 1e-9 / 1e-6 / 2e-5 of FP64;
  gels: rn <= 5e-7 gel units and um, vm, wm within 1e-9 in every storage
        (the gels feel the fluid only through the concentration exchange).

The batching workload: the 64 instances of workloads.batch_instances() at
full size (edges 35-97, 1-4 gels of 4^3-8^3, rates +-20%), FP32 and FP16,
10 steps in one Batch (HIP graphs, autotuned launch shape): every member
equals its solo run bit for bit, and the three smallest members (FP32) match
the FP64 oracle run on the host (u <= 1e-5, c <= 1e-6, rn <= 5e-7).

The serving workload: the first 16 jobs of workloads.poisson_jobs(rate=20)
(edges 41-161, 1-6 gels of 6^3, FP16) with horizons capped to 6-13 steps, on
a step clock of 0.05 s per step, so jobs arrive while others run: each job's
final state and every .dat file it wrote equal the job run alone.
"""
import hashlib
import os
import shutil
import sys

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

# relative L2 bounds at every checkpoint; with the reference's IBM ramp FP16
# keeps the 1e-3 velocity target on u; at full IBM stiffness from step 20 (a
# far stronger forcing: rho - 1 reaches 2e-2 near the gels) FP16 reaches
# 1.1e-3 at 101^3 and 1.5e-3 at 201^3
TOL = {
    "f32": {"u": 1e-5, "c": 3e-6, "rho": 1e-5},
    "f16": {"u": 1e-3, "c": 1e-4, "rho": 2e-3},
    "bf16": {"u": 1e-2, "c": 1e-3, "rho": 2e-2},
}
TOL_FULL_IBM = {"f16": {"u": 2e-3}, "bf16": {"u": 2e-2}}
# 201^3 with the reference's ramp: at step 50 the IBM stiffness is 0.25% of
# full, rms |u| over the sample is 1.2e-6 against rms (rho - 1) 1.5e-5, and
# FP32's accumulated rounding of the stored deviations (~2e-11 absolute
# after 1250 sub-steps) is 1.5e-5 of the velocity signal
TOL_CASE = {("walled201_four_gels_ref", "f32"): {"u": 2e-5}}


def say(*a):
    print(*a, flush=True)


def golden_errors(gold, n, gels, f):
    cells = gold["cells"]
    e = {"u": rel_l2(f.read("u")[cells], gold[f"u_{n}"]),
         "c": rel_l2(f.read("c")[cells], gold[f"c_{n}"]),
         "rho": rel_l2(f.read("rho")[cells], gold[f"rho_{n}"])}
    e["rn"] = max(float(np.abs(g.read("rn") - gold[f"gel{k}_rn_{n}"]).max()) for k, g in enumerate(gels))
    e["chem"] = max(rel_l2(g.read(x), gold[f"gel{k}_{x}_{n}"]) for k, g in enumerate(gels) for x in ("um", "vm", "wm"))
    cons, gc = f.conservation(), gold[f"conservation_{n}"]
    e["mass_per_cell"] = abs(cons[0] - gc[0]) / f.N
    # total momentum nearly cancels in a closed box: its error is stated on
    # the scale of the local momenta, sqrt(N) |u|_2 >= sum |rho u|
    e["momentum"] = float(np.linalg.norm(cons[1:] - gc[1:]) / (np.sqrt(f.N) * gold[f"norm_u_{n}"]))
    return e


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("case", ["walled101_two_gels_ref", "walled101_two_gels_full", "walled201_four_gels_ref",
                                  "walled201_four_gels_full"])
def test_large_lattice_vs_fp64_fingerprints(glsm, case, storage):
    import make_large_lattices as M

    path = os.path.join(HERE, "golden", case + ".npz")
    gold = np.load(path)
    cps = M.CASES[case]["checkpoints"]
    say(f"{case} {storage}: {cps[-1]} steps")
    gels, c, f = M.scene(glsm, case, storage=storage)
    try:
        assert tuple(gold["grid"]) == f.grid
        tol = dict(TOL[storage])
        if M.CASES[case]["ramp"] is not None:
            tol.update(TOL_FULL_IBM.get(storage, {}))
        tol.update(TOL_CASE.get((case, storage), {}))
        for it in range(cps[-1]):
            glsm.system_step(gels, c, f, it)
            if it + 1 in cps:
                e = golden_errors(gold, it + 1, gels, f)
                say(f"  {case} {storage} step {it + 1}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()))
                for k in ("u", "c", "rho"):
                    assert e[k] <= tol[k], (it + 1, k, e[k])
                assert e["rn"] <= 5e-7 and e["chem"] <= 1e-9
                assert e["mass_per_cell"] <= {"f32": 1e-9, "f16": 1e-6, "bf16": 2e-5}[storage]
        assert f.check_finite()["bad_cells"] == 0
    finally:
        for o in gels + [f, c]:
            o.close()


# ------------------------------------------------------------------ batching
def member_state(gels, f):
    out = {"u": f.read("u"), "c": f.read("c"), "fpost": f.read("fpost")}
    for k, g in enumerate(gels):
        for name in ("rn", "um", "vm", "wm", "un_norm", "Veln", "T0", "T2", "center_vm"):
            out[f"gel{k}.{name}"] = g.read(name)
    return out


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batch64_members_equal_solo_and_fp64(glsm, oracle, storage):
    G = glsm
    from glsm_fluid_3d_amd import workloads as W

    steps = 10
    specs = W.batch_instances(64)
    say(f"batch of {len(specs)} instances, {storage}, {sum(s.cells for s in specs)} cells, {steps} steps")
    b = G.Batch(storage)
    insts = [W.build(G, sp, storage) for sp in specs]
    try:
        for inst in insts:
            b.add(*inst)
        b.step(1)
        shape, _ = b.autotune(reps=2)
        say("autotuned launch", shape)
        b.step(steps - 1)
        G.synchronize()
        mixed = []
        for gels, c, f in insts:
            mixed.append(member_state(gels, f))
            assert f.check_finite()["bad_cells"] == 0
    finally:
        b.close()
        for gels, c, f in insts:
            for o in gels + [f, c]:
                o.close()
    for k, sp in enumerate(specs):
        gels, c, f = W.build(G, sp, storage)
        try:
            for it in range(steps):
                G.system_step(gels, c, f, it)
            alone = member_state(gels, f)
        finally:
            for o in gels + [f, c]:
                o.close()
        for name in alone:
            np.testing.assert_array_equal(mixed[k][name], alone[name], err_msg=f"member {k}: {name}")
    say(f"all {len(specs)} members equal their solo runs")
    if storage != "f32":
        return
    smallest = sorted(range(len(specs)), key=lambda i: (specs[i].cells, i))[:3]
    for k in smallest:
        sp = specs[k]
        ogels, oc, of = W.build(oracle, sp)
        for it in range(steps):
            oracle.system_step(ogels, oc, of, it)
        m = mixed[k]
        eu = rel_l2(m["u"], of.read("u"))
        ec = rel_l2(m["c"], of.read("c"))
        ern = max(float(np.abs(m[f"gel{j}.rn"] - og.read("rn")).max()) for j, og in enumerate(ogels))
        rates = [(round(g[2], 4), round(g[3], 4)) for g in sp.gels]
        say(f"  member {k} (edge {sp.edge}, rates {rates}) vs FP64: u {eu:.2e}, c {ec:.2e}, rn {ern:.1e}")
        assert eu <= 1e-5 and ec <= 1e-6 and ern <= 5e-7
        for j, og in enumerate(ogels):
            for name in ("um", "vm", "wm"):
                assert rel_l2(m[f"gel{j}.{name}"], og.read(name)) <= 1e-9, (k, j, name)


# ------------------------------------------------------------------ serving
def tree_digest(root):
    """{file name: sha256} of a job's output directory."""
    return {n: hashlib.sha256(open(os.path.join(root, n), "rb").read()).hexdigest() for n in sorted(os.listdir(root))}


def test_poisson_jobs_same_alone_and_in_the_mix(glsm, tmp_path, monkeypatch):
    G = glsm
    from glsm_fluid_3d_amd import serving as S
    from glsm_fluid_3d_amd import workloads as W

    jobs = W.poisson_jobs(seed=4, rate=20.0, duration=600.0, max_jobs=16)
    for k, j in enumerate(jobs):
        j.horizon = 6 + k % 8
    say(f"{len(jobs)} jobs, edges {[j.edge for j in jobs]}, {sum(j.cells for j in jobs)} cells")
    mixed = {}

    def keep(rec, gels, c, f):
        mixed[rec.job.job_id] = member_state(gels, f)

    root = tmp_path / "mix"
    srv = S.Server(G, storage="f16", out_root=str(root), max_chunk=3, on_complete=keep, writers=8)
    try:
        stats = srv.run(jobs, clock=S.StepClock(0.05))
    finally:
        srv.close()
    say({k: stats[k] for k in ("completed", "failed", "rejected", "batch_steps", "latency_p50_s", "latency_p99_s")})
    assert stats["completed"] == len(jobs) and not stats["failed"] and stats["output"]["error"] == ""
    recs = sorted(srv.done, key=lambda r: r.admitted)
    overlaps = sum(1 for r in recs for q in recs if r is not q and r.admitted < q.completed and q.admitted < r.completed)
    say(f"overlapping job pairs: {overlaps // 2}")
    assert overlaps > 0
    digests = {j.job_id: tree_digest(root / f"job{j.job_id}") for j in jobs}
    shutil.rmtree(root)
    for job in jobs:
        alone_dir = tmp_path / f"alone{job.job_id}"
        alone_dir.mkdir()
        monkeypatch.chdir(alone_dir)
        gels, c, f = W.build(G, job, "f16")
        try:
            for it in range(job.horizon):
                G.system_step(gels, c, f, it)
                for g in gels:
                    g.write_files(it)
                f.write_files(it)
            alone = member_state(gels, f)
        finally:
            for o in gels + [f, c]:
                o.close()  # joins the writer threads
        for k in alone:
            np.testing.assert_array_equal(mixed[job.job_id][k], alone[k], err_msg=f"job {job.job_id}: {k}")
        d = tree_digest(alone_dir)
        assert d and d == digests[job.job_id], f"job {job.job_id}: .dat files differ"
        monkeypatch.chdir(tmp_path)
        shutil.rmtree(alone_dir)
    say(f"all {len(jobs)} jobs equal alone and in the mix, {sum(len(d) for d in digests.values())} files")
