#!/usr/bin/env python3
"""Fragmented Gvoxels/s of the voxelize -> fracture path on one MI355X.

Workload (default): the mixed batch — 64 independent fracture jobs, each a
closed lathe vessel of 5k-300k triangles voxelized to a cubic grid with edge
uniform in [64, 384] (rounded to a multiple of 4, as the reference rounds its
voxelization sizes), 2-64 seeds drawn uniformly from the occupied cells,
distance uniform over {Euclidean, Manhattan, Chebyshev}, naive or flood 50/50
(flood with boundary noise amplitude 2), then boundary detection and the
per-fragment histogram.  RNG seed 4.  Meshes are uploaded to HBM before the
timed region.

One step = one pass of every job through the path.  `value` is the batched
form (every stage one launch over all 64 jobs, captured in HIP graphs):
sum of grid cells (n^3) over the jobs / wall time.  The same JSON line carries
the plain form (one job per launch), the served form (Poisson arrivals with
continuous batching: p50/p99 latency), the roofline of the dominant kernel
and the CPU oracle's rate on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "fragmented Gvoxels/s + p50/p99 job latency, batched/served vs plain, bit-exact"
BOX = ([-1.03, -1.03, -1.03], [1.03, 1.03, 1.03])
HBM_PEAK_GBS = 8000.0
# rocprofv3 summary of the default workload: FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a
# 1 GiB copy's counts, written as "traffic_bytes_per_launch <kernel> <bytes>"
TRAFFIC_SUMMARY = ROOT / "profiles" / "flood_relax_traffic_mixed64.txt"
DEFAULT_JOBS = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def job_specs(n, seed, emin=64, emax=384, smax=64, tmin=5000, tmax=300000):
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        e = int(rng.integers(emin, emax + 1)) // 4 * 4
        specs.append(dict(edge=max(e, 4), tris=int(rng.integers(tmin, tmax + 1)), seeds=int(rng.integers(2, smax + 1)),
                          dist=int(rng.integers(0, 3)), algo=int(rng.integers(0, 2)),
                          mesh_seed=int(rng.integers(1, 1 << 30)), seed_key=int(rng.integers(1, 1 << 40)),
                          noise_seed=int(rng.integers(1, 1 << 40))))
    return specs


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the default workload from the committed counter summary, or None."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    for line in TRAFFIC_SUMMARY.read_text().splitlines():
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def build_jobs(vf, specs, mc=False):
    meshes, jobs = [], []
    for s in specs:
        xyz, tri = vf.vessel_mesh(s["mesh_seed"], *vf.vessel_segments(s["tris"]))
        m = vf.DeviceMesh(xyz, tri)
        meshes.append(m)
        e = s["edge"]
        jobs.append(m.job((e, e, e), *BOX, s["seeds"], s["seed_key"], s["algo"], s["dist"],
                          noise_seed=s["noise_seed"], noise_amp=2 if s["algo"] == 1 else 0, boundary_size=1,
                          marching_cubes=1 if mc else 0))
    return meshes, jobs


# Algorithmic HBM bytes of each stage per step (the minimum traffic of the
# stage's definition, not of this implementation): cell stages count the
# bytes per cell they must read and write (labels 2 B, occupancy 1 B, flood
# keys 4 B); the voxelizer counts the triangle vertices (36 B per triangle);
# a flood relaxation round counts 4 KiB per brick visit (its 512 keys read
# and written once).
RELAX_BYTES_PER_VISIT = 512 * 4 * 2


def kernel_bytes(name, specs, results=None):
    V = [s["edge"] ** 3 for s in specs]
    T = [s["tris"] for s in specs]
    Vn = sum(v for v, s in zip(V, specs) if s["algo"] == 0)
    Vf = sum(v for v, s in zip(V, specs) if s["algo"] == 1)
    table = {
        "naive": 3 * Vn, "naive64": 3 * Vn,             # occupancy in, labels out
        "boundaries_hist": 4 * sum(V),                  # labels in and out
        "checksum": 2 * sum(V),                         # labels in
        "vox_fill": sum(V),                             # occupancy out
        "vox_setup": 36 * sum(T), "vox_shell": 36 * sum(T), "vox_parity": 36 * sum(T),
        "seed_count": sum(V),                           # occupancy in
        "flood_init": 5 * Vf,                           # occupancy in, keys out
        "flood_finish": 6 * Vf,                         # keys in, labels out
    }
    if name == "flood_relax" and results is not None:
        return RELAX_BYTES_PER_VISIT * sum(r["flood_bricks"] for r in results)
    return table.get(name)


def path_bytes(specs):
    """Algorithmic bytes of the whole path per step: voxelize 36 T + V, fracture 3 V, boundaries 4 V."""
    return sum(8 * s["edge"] ** 3 + 36 * s["tris"] for s in specs)


KERNELS = ["job_init", "vox_clear", "vox_setup", "scan", "vox_shell", "vox_parity", "vox_fill", "seed_count",
           "seed_pick", "naive", "naive64", "flood_init", "flood_seed", "flood_compact", "flood_relax", "flood_plan",
           "flood_finish", "boundaries_hist", "hist_summary", "checksum", "mc_count", "mc_emit", "mc_finish"]


def kernel_profile(vf, jobs, steps):
    """Per-kernel device time from HIP events recorded around every launch on
    the launch stream (graphs off, so each launch is timed)."""
    b = vf.Batch(graphs=False)
    res = b.run(jobs)
    vf.timing_reset()
    vf.timing_enable(True)
    for _ in range(steps):
        res = b.run(jobs)
    vf.timing_enable(False)
    prof = {}
    for k in KERNELS:
        ms, n = vf.timing_read(k)
        if n:
            prof[k] = {"total_ms": ms / steps, "launches": n / steps, "avg_ms": ms / n}
    b.close()
    return prof, res


def served(vf, specs, rate, n_req, budget, slots, seed):
    """Poisson arrivals at `rate` jobs/s; returns latencies and throughput."""
    meshes, jobs = build_jobs(vf, specs)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_req)
    order = rng.integers(0, len(jobs), size=n_req)
    srv = vf.Server(budget, slots)
    tickets, results = [None] * n_req, [None] * n_req
    done = threading.Event()

    def collect():
        for i in range(n_req):
            while tickets[i] is None:
                time.sleep(0.0002)
            results[i] = srv.wait(tickets[i])
        done.set()

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i in range(n_req):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[i] = srv.submit(jobs[order[i]])
    th.join()
    t1 = time.perf_counter()
    srv.close()
    lat = np.array([r["latency_ms"] for r in results])
    ok = all(r["status"] == 0 for r in results)
    vox = sum(specs[o]["edge"] ** 3 for o in order)
    out = {
        "requests": n_req, "rate_jobs_per_s": rate, "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
        "gvoxels_per_s": vox / (t1 - t0) / 1e9, "wall_s": t1 - t0, "all_ok": ok,
        "hbm_budget_gb": budget / 2**30, "slots": slots,
    }
    # each served result equals the same job run alone (bit-exact)
    sample = sorted(set(int(o) for o in order[:8]))
    alone = vf.Batch().run([jobs[o] for o in sample])
    first = {}
    for i, o in enumerate(order):
        first.setdefault(int(o), results[i]["checksum"])
    out["alone_equals_served"] = all(first[o] == a["checksum"] for o, a in zip(sample, alone))
    for m in meshes:
        m.close()
    return out


def cpu_baseline(specs, budget_s):
    """The CPU oracle (single thread) on the same jobs, in order, until the
    time budget is used: voxelize, seed, fracture, boundaries, histogram."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    from voxelfragmentml_amd import vfrag as vf

    done, vox, t_all = [], 0, 0.0
    for i, s in enumerate(specs):
        xyz, tri = vf.vessel_mesh(s["mesh_seed"], *vf.vessel_segments(s["tris"]))
        e = s["edge"]
        t0 = time.perf_counter()
        occ = O.voxelize(xyz, tri, *BOX, (e, e, e))
        g = O.occ_to_grid(occ)
        seeds = O.seed_uniform(g, s["seeds"], s["seed_key"])
        if s["algo"] == 0:
            lab = O.naive(g, seeds, s["dist"])
        else:
            lab = O.flood(g, seeds, 6 if s["dist"] == 1 else 26, s["noise_seed"], 2)
        b = O.detect_boundaries(lab, 1)
        O.histogram(b)
        t_all += time.perf_counter() - t0
        vox += e ** 3
        done.append(i)
        if t_all > budget_s:
            break
    return {"value": vox / t_all / 1e9, "unit": "Gvoxels/s", "cores": 1, "kind": "port",
            "sample": f"jobs {done[0]}-{done[-1]} of the mixed workload ({len(done)} jobs, {vox / 1e6:.1f} Mvoxels, "
                      f"{t_all:.1f} s), single-threaded C oracle, full path without marching cubes",
            "host_cpus": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=DEFAULT_JOBS)
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--skip", default="", help="comma list of: tune,plain,served,cpu,profile")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (a reference dispatch of known "
                         "traffic for rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = a.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "dispatches": 5}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); jobs are independent, so several GPUs would "
              f"each run their own share of the jobs, one process per GPU", file=sys.stderr)
        sys.exit(2)
    skip = set(x for x in a.skip.split(",") if x)
    from voxelfragmentml_amd import vfrag as vf

    if vf.device_count() < 1:
        raise SystemExit("no HIP device visible")
    specs = job_specs(a.jobs, 4)
    voxels = sum(s["edge"] ** 3 for s in specs)
    log(f"mixed workload: {len(specs)} jobs, {voxels / 1e9:.3f} Gvoxels per step")
    tuning = None
    if "tune" not in skip:
        # launch shapes from the autotuner (its own synthetic jobs per edge bucket), outside the timed region
        import tempfile
        t0 = time.perf_counter()
        with tempfile.TemporaryDirectory() as d:
            path = Path(d) / "tuning.txt"
            vf.tune_run(path, quick=True)
            tuning = {"table": vf.read_tuning(path), "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotuner: {tuning}")
    meshes, jobs = build_jobs(vf, specs)

    # ---- batched form (the headline)
    b = vf.Batch()
    for _ in range(a.warmup):
        ref = b.run(jobs)
    if a.warmup == 0:
        ref = None
    assert all(r["status"] == 0 for r in (ref or b.run(jobs))), "a job failed"
    ref = ref or b.run(jobs)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = b.run(jobs)
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    value = voxels / ((t1 - t0) / a.steps) / 1e9
    deterministic = [r["checksum"] for r in res] == [r["checksum"] for r in ref]
    log(f"batched: {ms_step:.2f} ms/step, {value:.2f} Gvoxels/s, deterministic={deterministic}")
    b.close()

    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gvoxels/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16 labels / u8 occupancy / int32 distances / fp32 SAT", "data": "synthetic",
        "config": {"workload": "mixed batch: 64 vessel jobs, edge U[64,384] (x4), 5k-300k triangles, 2-64 seeds, "
                               "distance uniform, naive/flood 50:50, flood noise 2, boundaries + histogram",
                   "jobs": len(specs), "voxels_per_step": voxels, "rng_seed": 4, "form": "batched (HIP graphs)"},
        "batched": {"gvoxels_per_s": value, "ms_per_step": ms_step, "deterministic": deterministic},
        "path_bytes": {"algorithmic_bytes_per_step": path_bytes(specs),
                       "achieved_gbs": round(path_bytes(specs) / (ms_step * 1e-3) / 1e9, 2),
                       "frac_of_hbm_peak": round(path_bytes(specs) / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "definition": "8 B per cell + 36 B per triangle per job"},
        "tuning": tuning,
    }

    if "plain" not in skip:
        plain = vf.plain_run(jobs)
        t0 = time.perf_counter()
        plain = vf.plain_run(jobs)
        t1 = time.perf_counter()
        pv = voxels / (t1 - t0) / 1e9
        same = [r["checksum"] for r in plain] == [r["checksum"] for r in ref]
        lat = np.array([r["latency_ms"] for r in plain])
        out["plain"] = {"gvoxels_per_s": pv, "ms_per_pass": (t1 - t0) * 1e3, "bit_exact_vs_batched": same,
                        "job_p50_ms": float(np.percentile(lat, 50)), "job_p99_ms": float(np.percentile(lat, 99))}
        out["batched_over_plain"] = value / pv
        log(f"plain: {pv:.2f} Gvoxels/s (batched/plain = {value / pv:.2f}x), bit-exact={same}")

    if "profile" not in skip:
        prof, pres = kernel_profile(vf, jobs, 2)
        out["kernels_ms_per_step"] = {k: round(v["total_ms"], 4) for k, v in prof.items()}
        cand = {k: v for k, v in prof.items() if kernel_bytes(k, specs, pres)}
        dom = max(cand, key=lambda k: cand[k]["total_ms"])
        per_launch = kernel_bytes(dom, specs, pres) / max(1.0, prof[dom]["launches"])
        achieved = per_launch / (prof[dom]["avg_ms"] * 1e-3) / 1e9
        out["roofline"] = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                           "bytes_per_launch": per_launch, "avg_launch_ms": prof[dom]["avg_ms"],
                           "launches_per_step": prof[dom]["launches"]}
        traffic = measured_traffic(dom) if a.jobs == DEFAULT_JOBS else None
        if traffic:
            out["roofline"]["traffic"] = traffic
            out["roofline"]["traffic_source"] = (f"{TRAFFIC_SUMMARY.relative_to(ROOT)}: rocprofv3 --pmc FETCH_SIZE and "
                                                 "WRITE_SIZE of this workload, each scaled by a 1 GiB copy's count")
        if dom == "flood_relax":
            visits = sum(r["flood_bricks"] for r in pres)
            out["roofline"]["unit_work"] = {"brick_visits_per_step": visits, "bytes_per_visit": RELAX_BYTES_PER_VISIT,
                                            "sweeps_per_step": sum(r["flood_sweeps"] for r in pres)}
        log(f"dominant kernel {dom}: {achieved:.1f} GB/s ({achieved / HBM_PEAK_GBS:.3f} of peak)")

    if "served" not in skip:
        sspecs = job_specs(48, 5, emax=512, smax=128, tmax=300000)
        out["served"] = served(vf, sspecs, a.serve_rate, a.serve_requests, 200 << 30, 64, 5)
        log(f"served: {out['served']}")

    if "cpu" not in skip:
        out["cpu_baseline"] = cpu_baseline(specs, a.cpu_seconds)
        log(f"cpu baseline: {out['cpu_baseline']}")

    for m in meshes:
        m.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
