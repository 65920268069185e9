#!/usr/bin/env python3
"""Benchmark of the gLSM_fluid_3D time step on one MI355X.

Headline workload (one "step" = one gel time step: every gel's
elasticity/chemistry, marker pack, 25 LBM sub-steps with immersed-boundary
coupling, the concentration step, repulsion and scatter -- sim.cpp's loop
body):
  201^3 D3Q19 lattice (physical box 100^3, h = 0.5, walls) with the
  concentration scalar, four 8x8x8-element gels at seeded non-overlapping
  centres, reference constants, chemistry at steady state + splitmix64
  perturbation of 0.01, FP32 arithmetic with FP32 deviation storage.  Both
  distribution buffers (1.23 GB) are far larger than the 256 MB Infinity
  Cache, so the stream/collide kernel runs from HBM.

Metric: MLUPS = lattice cells x 25 sub-steps x steps / wall seconds / 1e6,
inputs resident in HBM, no file output inside the timed region.

Printed JSON line (contract): metric, value, unit, n_gpus, steps, warmup,
ms_per_step, roofline (stream/collide kernel, live HIP-event timing on the
fluid's stream), cpu_baseline (the FP32 OpenMP restatement on a bounded
sample of the same workload), and extra lines: the 201^3 lattice in FP16 and
BF16 deviation storage, the 101^3 two-gel lattice (Infinity-Cache
resident), the 64-instance batch in FP32 and FP16, and the Poisson serving
stream (latency percentiles only once at least 20 jobs have completed).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
NSUB = 25


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cube101_two_gels(mod, storage=None, prec=None, seed=1):
    """101^3 lattice, two 6^3 gels placed as in sim.cpp (x = 1/3, 2/3)."""
    from glsm_fluid_3d_amd import workloads as W

    return W.two_gels_sim_cpp(mod, storage=storage, prec=prec, seed=seed)


def gel_array_3x3(mod, storage=None):
    """The 3 x 3 gel array of sim_array.cpp:31-72: a 140 x 140 x 10 box
    (281 x 281 x 21 nodes), nine 20 x 20 x 4-element gels at x, y in
    {30, 70, 110}, z = 7 -- 10,098 markers, whose O(M^2) repulsion is the
    large-M case of the coupler."""
    gels = [mod.Gel((20, 20, 4), (float(x), float(y), 7.0), gel_id=k + 1)
            for k, (y, x) in enumerate((y, x) for y in (30, 70, 110) for x in (30, 70, 110))]
    c = mod.Coupler(gels)
    f = mod.Fluid((140, 140, 10), c)
    if storage:
        f.set_storage(storage)
    return gels, c, f


def cilia_4x4(mod, storage=None):
    """The 4 x 4 cilia array of sim_cilia.cpp:36-84: a 100 x 100 x 30 box
    (201 x 201 x 61 nodes), sixteen 4 x 4 x 24-element gels at x, y in
    {10, 30, 50, 70}, z = 15, anchored at z = 3 -- every gel stepping, as
    once all of the driver's pulse delays have passed."""
    gels = []
    for k, (y, x) in enumerate((y, x) for y in (10, 30, 50, 70) for x in (10, 30, 50, 70)):
        g = mod.Gel((4, 4, 24), (float(x), float(y), 15.0), gel_id=k + 1)
        g.set_anchor_z(3.0)
        gels.append(g)
    c = mod.Coupler(gels)
    f = mod.Fluid((100, 100, 30), c)
    if storage:
        f.set_storage(storage)
    return gels, c, f


def cube201_four_gels(mod, storage):
    from glsm_fluid_3d_amd import workloads as W

    return W.four_gels_201(mod, storage=storage)


def run_batch(G, storage, steps, warmup, count=64, tune=True):
    from glsm_fluid_3d_amd import workloads as W

    insts = [W.build(G, sp, storage) for sp in W.batch_instances(count)]
    b = G.Batch(storage)
    for inst in insts:
        b.add(*inst)
    b.step(1)
    shape, tuned_ms = b.autotune(reps=3) if tune else (b.launch(), None)
    b.step(max(0, warmup - 1))
    b.enable_lbm_timing(10)
    G.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    G.synchronize()
    dt = time.perf_counter() - t0
    lbm_ms, samples = b.lbm_timing()
    cells = sum(f.N for _, _, f in insts)
    nsub = insts[0][2].nsub
    out = {"value": round(cells * nsub * steps / dt / 1e6, 1), "unit": "MLUPS", "steps": steps,
           "ms_per_step": round(dt / steps * 1e3, 4), "instances": count, "cells": cells,
           "gels": sum(len(g) for g, _, _ in insts), "markers": sum(c.M for _, c, _ in insts),
           "launch": shape, "hip_graph": True,
           "roofline": roofline(cells, 2 * 19 * G.STORAGE_BYTES[storage], lbm_ms)}
    per_shape = load_traffic(f"batch64_{storage}") if count == 64 else None
    # (an LDS reserve changes the work-groups per CU, not the bytes: the 201^3
    # FP32 launch moved 1,265.1 MB with and without one)
    key = f"nt{shape['nt']}_run{shape['run']}"
    out["roofline"]["traffic"] = (per_shape.get(key)
                                  if isinstance(per_shape, dict) and shape.get("grid") == 0 else None)
    out["roofline"]["kernel"] = "kb_lbm_step (batched persistent pull stream + collide)"
    out["roofline"]["samples"] = samples
    b.close()
    for gels, c, f in insts:
        for x in gels + [f, c]:
            x.close()
    return out


def run_slabs(G, storage="f32", parts=8, steps=40):
    """One GPU's share of a z-decomposed lattice: the 201^3 walled lattice
    split into `parts` z-slabs (25-26 planes + 2 halo planes each) stepped as
    a chain on this one GPU.  Reports the stream/collide launch at the slab's
    shape, the pack -> transfer -> unpack phase (transfer = a device copy
    standing in for the xGMI send/recv) and the face bytes one slab sends
    per sub-step and direction."""
    gels, c, f = cube201_four_gels(G, storage)
    chain = [G.Slab(f, z0, nz) for z0, nz in G.split_z(f.grid[2], parts)]
    try:
        for sl in chain:
            sl.load(f)
        G.slab_chain_step(chain, 5)
        G.synchronize()
        t0 = time.perf_counter()
        G.slab_chain_step(chain, steps)
        G.synchronize()
        dt = time.perf_counter() - t0
        for sl in chain:
            sl.enable_timing(True)
        G.slab_chain_step(chain, 10)
        tim = [sl.timing() for sl in chain]
        face = chain[0].face_bytes()
        full = [sl for sl in chain if sl.nz == max(x.nz for x in chain)]
        lbm_ms = float(np.median([tim[chain.index(sl)][0] for sl in full]))
        xch_ms = float(np.median([t[1] for t in tim]))
        nz = full[0].nz
        own = f.grid[0] * f.grid[1] * nz
        return {"slabs": parts, "slab_shape": f"{f.grid[0]}x{f.grid[1]}x{nz} (+2 halo planes)",
                "chain_MLUPS_one_gpu": round(f.N * steps / dt / 1e6, 1),
                "slab_lbm_ms": round(lbm_ms, 5),
                "slab_lbm_GBs_own_planes": round(own * 2 * 19 * G.STORAGE_BYTES[storage] / (lbm_ms * 1e-3) / 1e9, 1),
                "exchange_ms_chain_phase": round(xch_ms, 5), "face_bytes": face,
                "xgmi_us_per_face_at_153GBs": round(face / 153e9 * 1e6, 2)}
    finally:
        for sl in chain:
            sl.close()
        for x in gels + [c, f]:
            x.close()


def run_serve(G, seconds, outputs=True, seed=4, rate=20.0, drain=False, min_completed=None):
    """The serving workload: Poisson arrivals at `rate` jobs/s (20 by
    default), edge 41-161, 1-6 gels of 6^3, horizons 1000-20000 steps, FP16
    storage, .dat output streamed asynchronously (into a scratch directory
    removed afterwards).  Runs for `seconds` of wall time, or, with
    min_completed, until that many jobs have completed (at most `seconds`).
    With `drain`, arrivals stop after `seconds` and the server runs until
    every job has completed.  Latency percentiles are reported only over at
    least 20 completed jobs."""
    import shutil
    import tempfile

    from glsm_fluid_3d_amd import serving as S
    from glsm_fluid_3d_amd import workloads as W

    jobs = W.poisson_jobs(seed=seed, rate=rate, duration=seconds)
    root = tempfile.mkdtemp(prefix="glsm_serve_") if outputs else None
    srv = S.Server(G, storage="f16", out_root=root, keep_output=False, writers=8)
    try:
        stats = srv.run(jobs, duration=None if drain else seconds, progress_every=30.0,
                        until_completed=None if drain else min_completed)
    finally:
        srv.close()
        if root:
            shutil.rmtree(root, ignore_errors=True)
    arrived = [j for j in jobs if j.arrival <= stats["elapsed_s"]]
    stats["arrived"] = len(arrived)
    stats["offered_MLUPS"] = round(sum(j.cells * 25 * j.horizon for j in arrived) / stats["elapsed_s"] / 1e6, 1)
    if stats["completed"] < 20:
        for k in ("latency_p50_s", "latency_p99_s", "queue_wait_p50_s", "queue_wait_p99_s"):
            stats[k] = None
        stats["latency_note"] = f"only {stats['completed']} jobs completed: percentiles need at least 20"
    return stats


def run_gpu(G, gels, c, f, steps, warmup, every=10):
    for it in range(warmup):
        G.system_step(gels, c, f, it)
    f.enable_lbm_timing(every)
    G.synchronize()
    t0 = time.perf_counter()
    for it in range(warmup, warmup + steps):
        G.system_step(gels, c, f, it)
    G.synchronize()
    dt = time.perf_counter() - t0
    lbm_ms, samples = f.lbm_timing()
    return dt, lbm_ms, samples


def roofline(N, bytes_per_cell, lbm_ms, traffic=None):
    alg_bytes = N * bytes_per_cell
    achieved = alg_bytes / (lbm_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_lbm_step (fused pull stream + BGK/Guo collide)",
            "bytes_per_launch": alg_bytes, "mean_launch_ms": round(lbm_ms, 5)}


def load_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 counter summary."""
    p = os.path.join(ROOT, "profiles", "lbm_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        return json.load(open(p)).get(name)
    except (OSError, ValueError):
        return None


def host_cpu():
    """(model name, physical cores) of this host, from lscpu."""
    import subprocess

    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        cores = int(info.get("Core(s) per socket", "0")) * int(info.get("Socket(s)", "0"))
    except ValueError:
        cores = 0
    return info.get("Model name", "unknown"), cores or (os.cpu_count() or 1)


def cpu_baseline(target_s):
    """The FP32 OpenMP restatement (oracle/) on the headline workload: one
    untimed step, then as many timed steps as fit in about target_s
    seconds.  Threads: the CPU share this process may use ($OMP_NUM_THREADS,
    else all physical cores), bound one per core."""
    model, phys = host_cpu()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or phys
    threads = max(1, min(threads, phys))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle as O  # loads libgomp with the binding above

    O.set_threads(threads)
    gels, c, f = cube201_four_gels_oracle(O)
    N = f.N
    t0 = time.perf_counter()
    O.system_step(gels, c, f, 0)
    first = time.perf_counter() - t0
    steps = max(1, int(target_s / max(first, 1e-3)) - 1)
    t0 = time.perf_counter()
    for it in range(1, 1 + steps):
        O.system_step(gels, c, f, it)
    dt = time.perf_counter() - t0
    return {"value": round(N * NSUB * steps / dt / 1e6, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
            "cpu_model": model, "physical_cores": phys,
            "binding": f"OMP_PROC_BIND={os.environ['OMP_PROC_BIND']} OMP_PLACES={os.environ['OMP_PLACES']}",
            "sample": f"FP32 OpenMP restatement (oracle/), the same 201^3 four-gel workload, steps 1..{steps} after "
                      f"one untimed step ({N * NSUB * steps / 1e6:.0f} M lattice updates, {dt:.1f} s)"}


def cube201_four_gels_oracle(O):
    from glsm_fluid_3d_amd import workloads as W

    return W.four_gels_201(O, prec=32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--storage", default="f32", choices=["f32", "f32full", "f16", "bf16"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--extra-steps", type=int, default=400)
    ap.add_argument("--batch-steps", type=int, default=300)
    ap.add_argument("--serve-seconds", type=float, default=300.0,
                    help="serving window; the default run stops once 20 jobs have completed")
    ap.add_argument("--serve-completed", type=int, default=20,
                    help="stop the serving leg once this many jobs have completed (0 = run the whole window)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate of the serving workload")
    ap.add_argument("--serve-drain", action="store_true",
                    help="stop arrivals after --serve-seconds and run until every job has completed")
    ap.add_argument("--interp-mode", type=int, default=0, choices=[0, 1, 2, 3],
                    help="IBM interpolation form: 0 default (staged), 1 wavefront per marker, 2 thread per marker, "
                         "3 staged in LDS")
    ap.add_argument("--lds-reserve", type=int, default=-1,
                    help="LDS bytes reserved per stream/collide work-group (caps work-groups per CU); -1 automatic")
    ap.add_argument("--xcd-run", type=int, default=-1, choices=[-1, 0, 4, 16, 64],
                    help="XCD run length of the single-lattice stream/collide block order (-1 automatic)")
    ap.add_argument("--workload", default="all",
                    choices=["all", "101", "201-f32", "201-f16", "201-bf16", "copy", "batch-f32", "batch-f16",
                             "serve", "slab", "array", "cilia"],
                    help="profile one workload alone (no CPU baseline); 'copy' = the 1.23 GB calibration copy")
    args = ap.parse_args()
    if args.workload != "all":
        args.no_cpu_baseline = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}): the time step runs on a single "
              "device, and no multi-GPU path is implemented.", file=sys.stderr)
        sys.exit(2)

    import glsm_fluid_3d_amd as G

    G.init(0)
    G.set_interp_mode(args.interp_mode)
    G.set_lattice_xcd_run(args.xcd_run)
    G.set_lbm_lds_reserve(args.lds_reserve)
    if args.workload == "copy":
        nbytes = 2 * 19 * 4 * 8120601 // 2  # one 201^3 FP32 distribution buffer
        print(json.dumps({"workload": "plain 16-B copy", "bytes": nbytes,
                          "GBs": round(G.copy_bandwidth(nbytes, max(1, args.steps)), 1)}), flush=True)
        return
    if args.workload == "serve":
        r = run_serve(G, args.serve_seconds, rate=args.serve_rate, drain=args.serve_drain,
                      min_completed=args.serve_completed or None)
        r["workload"] = (f"Poisson {args.serve_rate:g} jobs/s for {args.serve_seconds:.0f} s"
                         f"{' then drained' if args.serve_drain else ''}, f16, async .dat output")
        print(json.dumps(r), flush=True)
        return
    if args.workload in ("array", "cilia"):
        gels, c, f = (gel_array_3x3 if args.workload == "array" else cilia_4x4)(G, args.storage)
        dt, ms, n = run_gpu(G, gels, c, f, args.steps, args.warmup)
        name = ("sim_array 3x3 gel array: 281x281x21 nodes, nine 20x20x4 gels" if args.workload == "array" else
                "sim_cilia 4x4 anchored gels: 201x201x61 nodes, sixteen 4x4x24 gels")
        print(json.dumps({"workload": name, "markers": c.M,
                          "value": round(f.N * f.nsub * args.steps / dt / 1e6, 1), "unit": "MLUPS",
                          "steps": args.steps, "ms_per_step": round(dt / args.steps * 1e3, 4),
                          "s_for_200000_steps": round(dt / args.steps * 200000, 1),
                          "roofline": roofline(f.N, 2 * 19 * G.STORAGE_BYTES[args.storage], ms)}), flush=True)
        for x in gels + [c, f]:
            x.close()
        return
    if args.workload == "slab":
        r = run_slabs(G, args.storage)
        r["workload"] = f"201^3 walled lattice in 8 z-slabs on one GPU, {args.storage}"
        print(json.dumps(r), flush=True)
        return
    if args.workload.startswith("batch"):
        st = args.workload.split("-")[1]
        r = run_batch(G, st, args.steps, args.warmup)
        r["workload"] = f"64 instances (edge 33-97, 1-4 gels of 4^3-8^3), {st}"
        print(json.dumps(r), flush=True)
        return
    if args.workload in ("101", "201-f32", "201-f16", "201-bf16"):
        st = args.storage if args.workload == "101" else args.workload.split("-")[1]
        gels, c, f = (cube101_two_gels(G, storage=st) if args.workload == "101" else cube201_four_gels(G, st))
        dt2, ms2, n2 = run_gpu(G, gels, c, f, args.steps, args.warmup)
        v2 = f.N * f.nsub * args.steps / dt2 / 1e6
        print(json.dumps({"workload": f"{args.workload} lattice, {st}", "value": round(v2, 1), "unit": "MLUPS",
                          "steps": args.steps, "ms_per_step": round(dt2 / args.steps * 1e3, 4),
                          "roofline": roofline(f.N, 2 * 19 * G.STORAGE_BYTES[st], ms2)}), flush=True)
        return

    # ---- headline: 201^3, four 8^3 gels, FP32 deviation storage, HBM-resident
    gels, c, f = cube201_four_gels(G, args.storage)
    N = f.N
    log(f"workload 201^3: N={N} nsub={f.nsub} markers={c.M} storage={args.storage}")
    dt, lbm_ms, samples = run_gpu(G, gels, c, f, args.steps, args.warmup, every=1)
    mlups = N * f.nsub * args.steps / dt / 1e6
    bpc = 2 * 19 * G.STORAGE_BYTES[args.storage]
    line = {
        "metric": "MLUPS (LBM lattice-node updates/s, full gel-fluid time step)",
        "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4), "gel_steps_per_s": round(args.steps / dt, 1),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "201^3 D3Q19 + concentration, four 8x8x8-element gels at seeded non-overlapping "
                               "centres, walls, seeded chemistry perturbation, 25 LBM sub-steps per gel step "
                               "(HBM-resident: 1.23 GB of distributions)",
                   "lattice": list(f.grid), "gels": len(gels), "markers": c.M, "nsub": f.nsub,
                   "storage": args.storage + (" (f - w_q deviations)" if args.storage != "f32full" else ""),
                   "parallelism": "single GPU"},
        "roofline": roofline(N, bpc, lbm_ms, load_traffic(f"cube201_four_gels_{args.storage}")),
    }
    line["roofline"]["samples"] = samples
    log(f"201^3: {mlups:.0f} MLUPS, {dt / args.steps * 1e3:.3f} ms/step, lbm kernel {lbm_ms * 1e3:.1f} us "
        f"({samples} samples)")
    for x in gels + [c, f]:
        x.close()
    copy_gbs = G.copy_bandwidth(19 * 4 * 8120601, 20)
    line["roofline"]["copy_GBs_1.23GB"] = round(copy_gbs, 1)
    log(f"plain copy of 1.23 GB: {copy_gbs:.0f} GB/s")

    if not args.no_extra:
        extra = {}
        for st in ("f16", "bf16"):
            gels, c, f = cube201_four_gels(G, st)
            dt2, ms2, n2 = run_gpu(G, gels, c, f, args.extra_steps, 20)
            v2 = f.N * f.nsub * args.extra_steps / dt2 / 1e6
            extra[f"201cube_4gels_{st}"] = {
                "value": round(v2, 1), "unit": "MLUPS", "steps": args.extra_steps,
                "ms_per_step": round(dt2 / args.extra_steps * 1e3, 4), "vs_headline": round(v2 / mlups, 3),
                "roofline": roofline(f.N, 2 * 19 * G.STORAGE_BYTES[st], ms2, load_traffic(f"cube201_four_gels_{st}"))}
            log(f"201^3 {st}: {v2:.0f} MLUPS ({v2 / mlups:.2f}x), lbm kernel {ms2 * 1e3:.1f} us")
            for x in gels + [c, f]:
                x.close()
        gels, c, f = cube101_two_gels(G, storage="f32")
        dt2, ms2, n2 = run_gpu(G, gels, c, f, 4 * args.extra_steps, 20)
        v2 = f.N * f.nsub * 4 * args.extra_steps / dt2 / 1e6
        extra["101cube_2gels_f32"] = {
            "value": round(v2, 1), "unit": "MLUPS", "steps": 4 * args.extra_steps,
            "ms_per_step": round(dt2 / (4 * args.extra_steps) * 1e3, 4),
            "note": "reference default scene (sim.cpp); both distribution buffers (157 MB) fit in the 256 MB "
                    "Infinity Cache, so its roofline fraction is cache-assisted, not an HBM figure",
            "roofline": roofline(f.N, 152, ms2, load_traffic("cube101_two_gels_f32"))}
        log(f"101^3 f32: {v2:.0f} MLUPS, lbm kernel {ms2 * 1e3:.1f} us")
        for x in gels + [c, f]:
            x.close()
        for st in ("f32", "f16"):
            r = run_batch(G, st, args.batch_steps, 20)
            r["vs_headline"] = round(r["value"] / mlups, 3)
            extra[f"batch64_{st}"] = r
            log(f"batch of 64 {st}: {r['value']:.0f} MLUPS ({r['vs_headline']:.2f}x), lbm kernel "
                f"{r['roofline']['mean_launch_ms'] * 1e3:.1f} us, launch {r['launch']}")
        r = run_serve(G, args.serve_seconds, min_completed=args.serve_completed or None)
        r["vs_headline"] = round(r["mlups"] / mlups, 3) if r.get("mlups") else None
        extra["serve_poisson20_f16"] = r
        log(f"serving {r['elapsed_s']:.0f} s: {r['completed']} done, p50 {r['latency_p50_s']} s, "
            f"p99 {r['latency_p99_s']} s, {r['mlups']:.0f} MLUPS")
        line["extra"] = extra

    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
