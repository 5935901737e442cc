#!/usr/bin/env python3
"""Fragmented Gvoxels/s and p50/p99 job latency of the voxelize -> fracture
path on one MI355X.

Headline workload: the mixed batch — 64 independent fracture jobs, each a
closed lathe vessel of 5k-300k triangles voxelized to a cubic grid with edge
uniform in [64, 384] (rounded to a multiple of 4, as the reference rounds its
voxelization sizes), 2-64 seeds drawn uniformly from the occupied cells,
distance uniform over {Euclidean, Manhattan, Chebyshev}, naive or flood 50/50
(flood with boundary noise amplitude 2), then boundary detection and the
per-fragment histogram.  RNG seed 4.  Meshes are uploaded to HBM before the
timed region.

One step = one pass of every job through the path.  `value` is the batched
form (every stage one launch over all 64 jobs, captured in HIP graphs):
sum of grid cells (n^3) over the jobs / wall time.  The same JSON line
carries the plain form (one job per launch), the served form (a Poisson
stream of 10k requests at 200 jobs/s, each drawn fresh from the mixed
distribution extended to 512^3 and 128 seeds, with continuous batching:
p50/p99 latency, and a sample checked against the oracle), the single-grid
configurations (128^3 naive, 256^3 flood, 512^3 naive + flood + marching
cubes) with per-stage times, the roofline of the dominant kernel and of
every stage, and the reference logic (the CPU oracle) timed on the host's
cores.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import json
import os
import platform
import statistics
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
# plain copy's counts, written as "traffic_bytes_per_launch <kernel> <bytes>"
TRAFFIC_SUMMARY = ROOT / "profiles" / "flood_relax_traffic_mixed64.txt"
DEFAULT_JOBS = 64
HOST_WORKERS = 16  # host threads the CPU legs use (the CPU share of one GPU on an 8-GPU, 128-core node)


def edge_bucket(edge):
    """The autotuner's bucket of a batch whose largest grid edge is edge."""
    return 0 if edge <= 64 else 1 if edge <= 128 else 2 if edge <= 256 else 3


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


def make_spec_job(vf, mesh, s, mc=False):
    e = s["edge"]
    return mesh.job((e, e, e), *BOX, s["seeds"], s["seed_key"], s["algo"], s["dist"],
                    noise_seed=s["noise_seed"], noise_amp=2 if s["algo"] == 1 else 0, boundary_size=1,
                    marching_cubes=1 if mc else 0)


def build_jobs(vf, specs, mc=False):
    meshes, jobs = [], []
    for s in specs:
        xyz, tri = vf.vessel_mesh(s["mesh_seed"], *vf.vessel_segments(s["tris"]))
        m = vf.DeviceMesh(xyz, tri)
        meshes.append(m)
        jobs.append(make_spec_job(vf, m, s, mc))
    return meshes, jobs


# ---------------------------------------------------------------- roofline ---
# Algorithmic HBM bytes per stage, the fixed definitions of the stage (not of
# this implementation): voxelize 36 B per triangle (FP32 vertices read) + 1 B
# per cell (occupancy written); naive or flood 3 B per cell (occupancy read,
# uint16 labels written); boundaries 4 B per cell (labels read and written);
# histogram 2 B per cell (labels read); marching cubes 2 B per cell + 12 B per
# vertex and per triangle.  A kernel that fuses stages counts each stage's bytes.
STAGES = {
    "voxelize": ["job_init", "vox_clear", "vox_setup", "scan", "vox_shell", "vox_parity", "vox_fill"],
    "seed": ["seed_count", "seed_pick"],
    "naive": ["naive", "naive64"],
    "flood": ["flood_init", "flood_seed", "flood_compact", "flood_relax", "flood_plan", "flood_finish"],
    "boundaries+histogram": ["boundaries_hist", "hist_summary"],
    "marching_cubes": ["mc_count", "mc_pairs", "mc_sort", "mc_emit", "mc_smooth"],
}
KERNELS = [k for ks in STAGES.values() for k in ks] + ["gather"]


def stage_bytes(specs, mc_out=None):
    V = [s["edge"] ** 3 for s in specs]
    Vn = sum(v for v, s in zip(V, specs) if s["algo"] == 0)
    Vf = sum(v for v, s in zip(V, specs) if s["algo"] == 1)
    out = {"voxelize": 36 * sum(s["tris"] for s in specs) + sum(V), "seed": 0, "naive": 3 * Vn, "flood": 3 * Vf,
           "boundaries+histogram": 4 * sum(V) + 2 * sum(V)}
    if mc_out is not None:
        out["marching_cubes"] = 2 * sum(V) + 12 * mc_out
    return out


def path_bytes(specs):
    """Per step: voxelize 36 T + V, fracture 3 V, boundaries 4 V, histogram 2 V."""
    return sum(10 * s["edge"] ** 3 + 36 * s["tris"] for s in specs)


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


class Records:
    """JSON-lines records, one per job (dims, seeds, stage times, results), then a run summary; no-op without a
    path."""

    def __init__(self, path):
        self.f = open(path, "w") if path else None

    def __bool__(self):
        return self.f is not None

    def emit(self, rec):
        if self.f:
            self.f.write(json.dumps(rec) + "\n")

    def close(self):
        if self.f:
            self.f.close()
            self.f = None


def spec_record(s):
    e = s["edge"]
    return {"dims": [e, e, e], "triangles": s["tris"], "seeds": s["seeds"], "seed_key": s["seed_key"],
            "algorithm": "naive" if s["algo"] == 0 else "flood", "distance": ["euclidean", "manhattan", "chebyshev"][s["dist"]],
            "noise_seed": s["noise_seed"] if s["algo"] == 1 else None}


RESULT_KEYS = ["status", "occupied", "num_labels", "flood_levels", "flood_bricks", "flood_sweeps", "checksum",
               "mc_vertices", "mc_triangles"]


def job_stage_times(vf, job):
    """One job alone (graphs off): device ms per stage from HIP events around every launch."""
    b = vf.Batch(graphs=False)
    b.run([job])
    vf.timing_reset()
    vf.timing_enable(True)
    r = b.run([job])[0]
    vf.timing_enable(False)
    b.close()
    stage = {}
    for st, ks in STAGES.items():
        ms = sum(vf.timing_read(k)[0] for k in ks)
        if ms:
            stage[st] = round(ms, 4)
    return stage, r


def roofline(prof, specs, res, traffic_ok):
    sb = stage_bytes(specs)
    stages = {}
    for st, ks in STAGES.items():
        ms = sum(prof[k]["total_ms"] for k in ks if k in prof)
        if ms <= 0 or st not in sb:
            continue
        gbs = sb[st] / (ms * 1e-3) / 1e9
        stages[st] = {"ms_per_step": round(ms, 4), "algorithmic_bytes": sb[st], "achieved_gbs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dom = max((k for k in prof if k != "gather"), key=lambda k: prof[k]["total_ms"])
    st = next(s for s, ks in STAGES.items() if dom in ks)
    launches = prof[dom]["launches"]
    per_launch = sb[st] / launches  # the stage's bytes spread over the dominant kernel's launches
    achieved = per_launch / (prof[dom]["avg_ms"] * 1e-3) / 1e9
    out = {"kernel": dom, "stage": st, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
           "bytes_per_launch": per_launch, "avg_launch_ms": prof[dom]["avg_ms"], "launches_per_step": launches,
           "definition": f"{st} stage bytes per step ({'3 B per flood cell' if st == 'flood' else 'per-stage rule'}) "
                         f"/ launches per step / average launch time (HIP events on the launch stream)"}
    if dom == "flood_relax":
        visits = sum(r["flood_bricks"] for r in res)
        out["limiter"] = ("latency of each visit's dependent chain (halo loads, then sweeps of LDS round trips) over "
                          "dependent rounds: 13 of 18 waves per CU resident (4 x SQ_WAVE_CYCLES / SQ_BUSY_CYCLES per "
                          "CU), waiting on results 42 % and on issue 30 % of their time (SQ_WAIT_ANY, SQ_WAIT_INST_ANY), "
                          "VALU at 40 % of the rate k_vox_shell sustains (SQ_INSTS_VALU); not HBM")
        out["unit_work"] = {"brick_visits_per_step": visits, "visit_bytes_per_step": 4096 * visits,
                            "sweeps_per_step": sum(r["flood_sweeps"] for r in res),
                            "visit_gbs": round(4096 * visits / launches / (prof[dom]["avg_ms"] * 1e-3) / 1e9, 1)}
    traffic = measured_traffic(dom) if traffic_ok else None
    if traffic:
        out["traffic"] = traffic
        out["traffic_source"] = (f"{TRAFFIC_SUMMARY.relative_to(ROOT)}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                 "of this workload, each scaled by a plain copy's count")
    return out, stages


# ------------------------------------------------------------------ served ---
def served(vf, rate, n_req, budget, slots, seed, pool_size, check_n, records=None):
    """Poisson arrivals at `rate` jobs/s.  Each request is a fresh draw from the mixed distribution extended
    to 512^3 and 128 seeds (edge, seeds, distance, algorithm, seed and noise keys), on a mesh taken from a pool
    of `pool_size` device-resident vessels of 5k-300k triangles (uploaded before the stream starts)."""
    pool_specs = job_specs(pool_size, seed + 1000)
    pool, host = [], []
    for s in pool_specs:
        xyz, tri = vf.vessel_mesh(s["mesh_seed"], *vf.vessel_segments(s["tris"]))
        pool.append(vf.DeviceMesh(xyz, tri))
        host.append((xyz, tri))
    rng = np.random.default_rng(seed)
    reqs = job_specs(n_req, seed, emax=512, smax=128)
    which = rng.integers(0, pool_size, size=n_req)
    for s, w in zip(reqs, which):
        s["tris"] = pool_specs[w]["tris"]
    jobs = [make_spec_job(vf, pool[w], s, mc=True) for s, w in zip(reqs, which)]
    gaps = rng.exponential(1.0 / rate, size=n_req)
    srv = vf.Server(budget, slots)
    tickets, results = [None] * n_req, [None] * n_req

    def collect():
        for i in range(n_req):
            while tickets[i] is None:
                time.sleep(0.0002)
            results[i] = srv.wait(tickets[i])

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i in range(n_req):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets[i] = srv.submit(jobs[i])
        if (i + 1) % 2000 == 0:
            log(f"served: {i + 1} of {n_req} requests submitted")
    th.join()
    t1 = time.perf_counter()
    mem = srv.memory()
    srv.close()
    lat = np.array([r["latency_ms"] for r in results])
    vox = sum(s["edge"] ** 3 for s in reqs)
    if records:
        for i, (s, r) in enumerate(zip(reqs, results)):
            records.emit({"form": "served", "request": i, **spec_record(s), "latency_ms": round(r["latency_ms"], 4),
                          "queue_ms": round(r["queue_ms"], 4), **{k: r[k] for k in RESULT_KEYS}})
    out = {
        "requests": n_req, "rate_jobs_per_s": rate, "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
        "gvoxels_per_s": vox / (t1 - t0) / 1e9, "wall_s": t1 - t0, "all_ok": all(r["status"] == 0 for r in results),
        "hbm_budget_gb": budget / 2**30, "slots": slots, "mesh_pool": pool_size,
        "distribution": "edge U[64,512] (x4), seeds U[2,128], distance uniform, naive/flood 50:50, flood noise 2, "
                        "boundaries + histogram + marching cubes of every fragment",
        "mc_triangles": sum(r["mc_triangles"] for r in results),
        "max_unreserved_bytes": mem["max_unreserved_bytes"],
    }
    # a sample of requests: the served result equals the job run alone and the oracle
    pick = sorted(rng.choice(n_req, size=min(check_n, n_req), replace=False).tolist())
    alone = vf.Batch().run([jobs[i] for i in pick])
    out["alone_equals_served"] = all(a["checksum"] == results[i]["checksum"] for a, i in zip(alone, pick))
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O

    def ref(i):
        s = reqs[i]
        xyz, tri = host[which[i]]
        e = s["edge"]
        g = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, (e, e, e)))
        seeds = O.seed_uniform(g, s["seeds"], s["seed_key"])
        lab = O.naive(g, seeds, s["dist"]) if s["algo"] == 0 else O.flood(g, seeds, 6 if s["dist"] == 1 else 26,
                                                                         s["noise_seed"], 2)
        return O.checksum(O.detect_boundaries(lab, 1))

    t2 = time.perf_counter()
    with cf.ThreadPoolExecutor(HOST_WORKERS) as ex:
        refs = list(ex.map(ref, pick))
    out["oracle_sample"] = {"jobs": len(pick), "equal": sum(int(r == results[i]["checksum"]) for r, i in zip(refs, pick)),
                            "seconds": round(time.perf_counter() - t2, 1)}
    for m in pool:
        m.close()
    return out


# ---------------------------------------------------------- configurations ---
CONFIGS = [
    dict(name="128^3 naive, 8 seeds, Manhattan (~20k-triangle vessel)", edge=128, tris=20000, mesh_seed=1, seeds=8,
         seed_key=1, dist=1, algos=[0], mc=False),
    dict(name="256^3 flood, 32 seeds, Euclidean -> 26 neighbours, noise 2, histogram (~200k triangles)", edge=256,
         tris=200000, mesh_seed=2, seeds=32, seed_key=2, dist=0, algos=[1], mc=False),
    dict(name="512^3 naive and flood, 128 seeds, Chebyshev, marching cubes (~1M triangles)", edge=512,
         tris=1000000, mesh_seed=3, seeds=128, seed_key=3, dist=2, algos=[0, 1], mc=True),
]


def config_runs(vf, reps):
    """Single-grid form of each configuration: every call synchronises, so the wall time of each call is its
    stage time.  Median of `reps` runs per stage."""
    out = []
    for c in CONFIGS:
        e = c["edge"]
        dims = (e, e, e)
        xyz, tri = vf.vessel_mesh(c["mesh_seed"], *vf.vessel_segments(c["tris"]))
        mesh = vf.DeviceMesh(xyz, tri)
        times = {}

        def clock(name, fn):
            t0 = time.perf_counter()
            r = fn()
            times.setdefault(name, []).append((time.perf_counter() - t0) * 1e3)
            return r

        mc_info = None
        with vf.Grid(dims, *BOX) as g:
            for rep in range(reps + 1):
                g.set_aabb(*BOX, dims)  # RegularGrid::setAABB: an empty grid, so the voxelizer writes it whole
                clock("voxelize", lambda: g.voxelize_device(mesh))
                base = g.read()
                seeds = clock("seed", lambda: g.seed_uniform(c["seeds"], c["seed_key"]))
                for a in c["algos"]:
                    if a != c["algos"][0]:
                        g.write(base)
                    if a == 0:
                        clock("naive", lambda: g.fracture_naive(seeds, c["dist"]))
                    else:
                        clock("flood", lambda: g.fracture_flood(seeds, c["dist"], noise_seed=c["seed_key"],
                                                                noise_amp=2))
                clock("boundaries", lambda: g.detect_boundaries(1))
                clock("histogram", lambda: g.histogram())
                if c["mc"]:
                    mc_info = clock("marching_cubes", lambda: g.marching_cubes_info())
                if rep == 0:
                    times.clear()  # the first pass warms up
        med = {k: statistics.median(v) for k, v in times.items()}
        total = sum(med.values())
        V = e ** 3
        occ = int((base != 0).sum())
        # batched form of the same job (flood when the configuration runs it)
        job = mesh.job(dims, *BOX, c["seeds"], c["seed_key"], c["algos"][-1], c["dist"], noise_seed=c["seed_key"],
                       noise_amp=2 if c["algos"][-1] == 1 else 0, boundary_size=1, marching_cubes=1 if c["mc"] else 0)
        bt = []
        with vf.Batch() as b:
            b.run([job])
            for _ in range(reps):
                t0 = time.perf_counter()
                b.run([job])
                bt.append((time.perf_counter() - t0) * 1e3)
        entry = {"config": c["name"], "voxels": V, "occupied": occ, "stage_ms": {k: round(v, 3) for k, v in med.items()},
                 "plain_gvoxels_per_s": round(V * len(c["algos"]) / (total * 1e-3) / 1e9, 3),
                 "plain_occupied_gvoxels_per_s": round(occ * len(c["algos"]) / (total * 1e-3) / 1e9, 4),
                 "batched_ms": round(statistics.median(bt), 3),
                 "batched_gvoxels_per_s": round(V / (statistics.median(bt) * 1e-3) / 1e9, 3)}
        if mc_info:
            entry["mc_triangles"] = mc_info["num_triangles"]
            entry["mc_vertices"] = mc_info["num_vertices"]
            entry["mc_triangles_per_s"] = round(mc_info["num_triangles"] / (med["marching_cubes"] * 1e-3), 1)
        out.append(entry)
        log(f"config {c['name']}: {entry}")
        mesh.close()
    return out


# -------------------------------------------------------------- CPU leg ---
def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(specs, n_jobs, runs, single_jobs):
    """The oracle (the reference logic restated in C) on the host's cores: jobs of the mixed workload spread
    over a pool of threads (the oracle's C calls release the interpreter lock), median wall time of `runs`
    passes over the same jobs, per-stage CPU seconds; plus one single-threaded pass over fewer jobs."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle as O
    from voxelfragmentml_amd import vfrag as vf

    sample = specs[:n_jobs]
    meshes = [vf.vessel_mesh(s["mesh_seed"], *vf.vessel_segments(s["tris"])) for s in sample]
    lock = threading.Lock()

    def one(i, stages):
        s, (xyz, tri) = sample[i], meshes[i]
        e = s["edge"]
        t = [time.perf_counter()]
        occ = O.voxelize(xyz, tri, *BOX, (e, e, e))
        t.append(time.perf_counter())
        g = O.occ_to_grid(occ)
        seeds = O.seed_uniform(g, s["seeds"], s["seed_key"])
        t.append(time.perf_counter())
        if s["algo"] == 0:
            lab = O.naive(g, seeds, s["dist"])
        else:
            lab = O.flood(g, seeds, 6 if s["dist"] == 1 else 26, s["noise_seed"], 2)
        t.append(time.perf_counter())
        b = O.detect_boundaries(lab, 1)
        t.append(time.perf_counter())
        O.histogram(b)
        t.append(time.perf_counter())
        names = ["voxelize", "seed", "naive" if s["algo"] == 0 else "flood", "boundaries", "histogram"]
        with lock:
            for k, a, c in zip(names, t, t[1:]):
                stages[k] = stages.get(k, 0.0) + (c - a)

    workers = min(HOST_WORKERS, len(os.sched_getaffinity(0)))
    walls, stages = [], {}
    for r in range(runs):
        st = {}
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(workers) as ex:
            list(ex.map(lambda i: one(i, st), range(len(sample))))
        walls.append(time.perf_counter() - t0)
        stages = st
        log(f"cpu baseline: pass {r + 1} of {runs}: {walls[-1]:.2f} s")
    vox = sum(s["edge"] ** 3 for s in sample)
    st1 = {}
    t0 = time.perf_counter()
    for i in range(single_jobs):
        one(i, st1)
    t1 = time.perf_counter() - t0
    vox1 = sum(s["edge"] ** 3 for s in sample[:single_jobs])
    return {"value": vox / statistics.median(walls) / 1e9, "unit": "Gvoxels/s", "cores": workers, "kind": "port",
            "sample": f"jobs 0-{len(sample) - 1} of the mixed workload ({len(sample)} jobs, {vox / 1e6:.1f} Mvoxels), "
                      f"oracle C restatement (gcc -O2) on {workers} threads (one job per thread at a time), median "
                      f"of {runs} passes; voxelize, seed, fracture, boundaries, histogram",
            "pass_seconds": [round(w, 3) for w in walls],
            "stage_cpu_seconds": {k: round(v, 3) for k, v in stages.items()},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "single_thread": {"value": vox1 / t1 / 1e9, "jobs": single_jobs, "seconds": round(t1, 2),
                              "stage_cpu_seconds": {k: round(v, 3) for k, v in st1.items()}}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--jobs", type=int, default=DEFAULT_JOBS)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-check", type=int, default=32, help="served requests compared with the oracle")
    ap.add_argument("--cpu-jobs", type=int, default=32)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--skip", default="", help="comma list of: tune,extract,plain,served,cpu,profile,configs")
    ap.add_argument("--records", default="", metavar="PATH",
                    help="write JSON-lines records: one per batched job (dims, seeds, per-stage device ms of the job "
                         "run alone, results), one per served request (latency, queue time), then the run summary")
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
    meshes, jobs = build_jobs(vf, specs)
    # the same batch with the extract stage: every job's fragment surfaces by marching cubes
    mjobs = [make_spec_job(vf, m, s, mc=True) for m, s in zip(meshes, specs)] if "extract" not in skip else None
    if "tune" not in skip:
        # launch shapes from the autotuner, tuned on this batch with its extract stage (every stage the bench
        # times), outside the timed region: the batch's median time with the default shapes and the tuned ones
        import tempfile
        t0 = time.perf_counter()
        with tempfile.TemporaryDirectory() as d:
            path = Path(d) / "tuning.txt"
            before, after = vf.tune_jobs(mjobs or jobs, path, quick=True)
            table = vf.read_tuning(path)
        tuning = {"tuned_on": "the batch with marching cubes" if mjobs else "the batch",
                  "default_ms_per_step": round(before, 3), "tuned_ms_per_step": round(after, 3),
                  "shape": table[edge_bucket(max(sp["edge"] for sp in specs))],
                  "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotuner: {tuning}")

    # ---- batched form (the headline)
    b = vf.Batch()
    ref = None
    for _ in range(max(1, a.warmup)):
        ref = b.run(jobs)
    assert all(r["status"] == 0 for r in ref), "a job failed"
    t0 = time.perf_counter()
    for _ in range(a.steps):
        res = b.run(jobs)
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / a.steps
    value = voxels / ((t1 - t0) / a.steps) / 1e9
    deterministic = [r["checksum"] for r in res] == [r["checksum"] for r in ref]
    occupied = sum(r["occupied"] for r in res)
    log(f"batched: {ms_step:.2f} ms/step, {value:.2f} Gvoxels/s, deterministic={deterministic}")
    b.close()

    pb = path_bytes(specs)
    out = {
        "metric": METRIC, "value": round(value, 4), "unit": "Gvoxels/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16 labels / u8 occupancy / int32 distances / fp32 SAT", "data": "synthetic",
        "config": {"workload": "mixed batch: 64 vessel jobs, edge U[64,384] (x4), 5k-300k triangles, 2-64 seeds, "
                               "distance uniform, naive/flood 50:50, flood noise 2, boundaries + histogram",
                   "jobs": len(specs), "voxels_per_step": voxels, "rng_seed": 4, "form": "batched (HIP graphs)"},
        "batched": {"gvoxels_per_s": value, "ms_per_step": ms_step, "deterministic": deterministic,
                    "occupied_gvoxels_per_s": occupied / (ms_step * 1e-3) / 1e9},
        "path_bytes": {"algorithmic_bytes_per_step": pb, "achieved_gbs": round(pb / (ms_step * 1e-3) / 1e9, 2),
                       "frac_of_hbm_peak": round(pb / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "definition": "per job: voxelize 36 T + V, fracture 3 V, boundaries 4 V, histogram 2 V"},
        "tuning": tuning,
    }

    if "extract" not in skip:
        # the extract stage: all jobs' marching cubes in one launch set after the fracture, grids resident
        bx = vf.Batch()
        xr = bx.run(mjobs)
        assert all(r["status"] == 0 for r in xr), "a job failed"
        t0 = time.perf_counter()
        for _ in range(a.steps):
            xr = bx.run(mjobs)
        t1 = time.perf_counter()
        bx.close()
        xms = (t1 - t0) * 1e3 / a.steps
        tris = sum(r["mc_triangles"] for r in xr)
        out["batched_extract"] = {"gvoxels_per_s": voxels / (xms * 1e-3) / 1e9, "ms_per_step": xms,
                                  "mc_triangles_per_step": tris, "mc_vertices_per_step": sum(r["mc_vertices"] for r in xr),
                                  "extract_ms_per_step": xms - ms_step,
                                  "mc_triangles_per_s": tris / ((xms - ms_step) * 1e-3) if xms > ms_step else None,
                                  "same_labels_as_batched": [r["checksum"] for r in xr] == [r["checksum"] for r in ref]}
        log(f"batched with extract: {out['batched_extract']}")

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
        out["roofline"], out["stage_roofline"] = roofline(prof, specs, pres, a.jobs == DEFAULT_JOBS)
        if "extract" not in skip:
            # the extract stage's own roofline, from the same batch with marching cubes
            mprof, mres = kernel_profile(vf, mjobs, 2)
            mc_out = sum(r["mc_vertices"] + r["mc_triangles"] for r in mres)
            ms = sum(mprof[k]["total_ms"] for k in STAGES["marching_cubes"] if k in mprof)
            mb = stage_bytes(specs, mc_out)["marching_cubes"]
            out["stage_roofline"]["marching_cubes"] = {
                "ms_per_step": round(ms, 4), "algorithmic_bytes": mb, "achieved_gbs": round(mb / (ms * 1e-3) / 1e9, 1),
                "frac": round(mb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "kernels_ms": {k: round(mprof[k]["total_ms"], 4) for k in STAGES["marching_cubes"] if k in mprof}}
        log(f"roofline: {out['roofline']}")
        log(f"stages: {out['stage_roofline']}")

    records = Records(a.records)
    if records:
        # per-job records: every job of the batch alone with events on (outside every timed region)
        xres = xr if "extract" not in skip else None
        for i, (s, job, r) in enumerate(zip(specs, mjobs or jobs, res)):
            stage, alone = job_stage_times(vf, job)
            rec = {"form": "batched", "job": i, **spec_record(s), "voxels": s["edge"] ** 3,
                   **{k: r[k] for k in RESULT_KEYS}, "stage_ms_alone": stage,
                   "extract": bool(mjobs), "alone_equals_batched": alone["checksum"] == r["checksum"]}
            if xres is not None:
                rec["mc_vertices"], rec["mc_triangles"] = xres[i]["mc_vertices"], xres[i]["mc_triangles"]
            records.emit(rec)
        log(f"records: {len(specs)} batched jobs written to {a.records}")

    if "configs" not in skip:
        out["configs"] = config_runs(vf, 3)

    if "served" not in skip:
        out["served"] = served(vf, a.serve_rate, a.serve_requests, 200 << 30, 64, 5, 64, a.serve_check, records)
        log(f"served: {out['served']}")

    if "cpu" not in skip:
        out["cpu_baseline"] = cpu_baseline(specs, a.cpu_jobs, a.cpu_runs, 4)
        log(f"cpu baseline: {out['cpu_baseline']}")

    for m in meshes:
        m.close()
    if records:
        summary = {"form": "summary", "gvoxels_per_s": out["value"], "ms_per_step": out["ms_per_step"],
                   "hbm_frac": out["path_bytes"]["frac_of_hbm_peak"]}
        if "served" in out:
            summary.update(served_p50_ms=out["served"]["p50_ms"], served_p99_ms=out["served"]["p99_ms"])
        if "plain" in out:
            summary.update(plain_job_p50_ms=out["plain"]["job_p50_ms"], plain_job_p99_ms=out["plain"]["job_p99_ms"])
        records.emit(summary)
        records.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
