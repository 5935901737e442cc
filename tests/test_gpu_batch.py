"""Batched, plain and served forms: each job's result (grid checksum,
histogram, occupied count, labels) must equal the single-grid path and the
oracle, alone and in a mix of uneven jobs; failing jobs must not disturb the
others and must release their slot and memory."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BOX = ([-1.03] * 3, [1.03] * 3)


def mixed_jobs(vf, n, seed, keep, edges=(40, 96), mc_every=0, locations=None):
    rng = np.random.default_rng(seed)
    jobs, meta = [], []
    for i in range(n):
        e = [int(rng.integers(edges[0], edges[1] + 1)) for _ in range(3)]
        tris = int(rng.integers(3000, 12000))
        xyz, tri = vf.vessel_mesh(int(rng.integers(1, 1 << 30)), *vf.vessel_segments(tris))
        algo = int(rng.integers(0, 2))
        dist = int(rng.integers(0, 3))
        S = int(rng.integers(2, 25))
        amp = int(rng.integers(0, 3)) if algo == 1 else 0
        labels = np.zeros(e, np.uint16)
        hist = np.zeros(65536, np.uint32)
        keep.extend([labels, hist])
        loc = locations[i % len(locations)] if locations else vf.SEED_BOTH
        j = vf.make_job(xyz, tri, e, *BOX, S, int(rng.integers(1, 1 << 40)), algo, dist,
                        noise_seed=int(rng.integers(1, 1 << 40)), noise_amp=amp,
                        boundary_size=1, marching_cubes=1 if (mc_every and i % mc_every == 0) else 0,
                        seed_location=loc, keep=keep)
        j.labels_out = labels.ctypes.data
        j.histogram_out = hist.ctypes.data
        jobs.append(j)
        meta.append(dict(xyz=xyz, tri=tri, dims=e, S=S, algo=algo, dist=dist, amp=amp, labels=labels, hist=hist,
                         seed=j.seed, noise_seed=j.noise_seed, location=loc))
    return jobs, meta


def oracle_job(m):
    occ = O.voxelize(m["xyz"], m["tri"], *BOX, m["dims"])
    g = O.occ_to_grid(occ)
    seeds = O.seed_uniform(g, m["S"], m["seed"], m.get("location", 2))
    if m["algo"] == 0:
        lab = O.naive(g, seeds, m["dist"])
    else:
        lab = O.flood(g, seeds, 6 if m["dist"] == 1 else 26, m["noise_seed"], m["amp"])
    return O.detect_boundaries(lab, 1), int(occ.sum())


def test_batch_matches_oracle_and_plain(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 10, 1, keep)
    b = vf.Batch()
    res = b.run(jobs)
    for m, r in zip(meta, res):
        assert r["status"] == 0, r
        ref, occ = oracle_job(m)
        np.testing.assert_array_equal(m["labels"], ref)
        np.testing.assert_array_equal(m["hist"], O.histogram(ref))
        assert r["occupied"] == occ
        assert r["checksum"] == O.checksum(ref)
        assert r["num_labels"] == int(np.count_nonzero(O.histogram(ref)[2:]))
    # replay (graphs) gives identical results
    res2 = b.run(jobs)
    assert [r["checksum"] for r in res2] == [r["checksum"] for r in res]
    # plain form, one job per launch, same answers
    plain = vf.plain_run(jobs)
    assert [r["checksum"] for r in plain] == [r["checksum"] for r in res]
    # without graphs
    b2 = vf.Batch(graphs=False)
    assert [r["checksum"] for r in b2.run(jobs)] == [r["checksum"] for r in res]


def test_batch_seed_locations_match_oracle(vf):
    # jobs seeding anywhere get their candidate counts from the voxelizer's fill, the others from
    # the candidate-count stage: a mix of both in one batch
    keep = []
    jobs, meta = mixed_jobs(vf, 6, 21, keep, edges=(40, 72), locations=[vf.SEED_OUTER, vf.SEED_BOTH, vf.SEED_INNER])
    b = vf.Batch()
    res = b.run(jobs)
    b.close()
    for r, m in zip(res, meta):
        print("seed location", m["location"], m["dims"], m["S"], flush=True)
        assert r["status"] == 0, r
        ref, occ = oracle_job(m)
        np.testing.assert_array_equal(m["labels"], ref)
        assert r["occupied"] == occ


def test_batch_alone_equals_in_mix(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 12, 7, keep)
    b = vf.Batch()
    mix = b.run(jobs)
    for i in (0, 5, 11):
        alone = vf.Batch().run([jobs[i]])[0]
        assert alone["checksum"] == mix[i]["checksum"] and alone["occupied"] == mix[i]["occupied"]


def test_batch_failures_isolated(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 6, 3, keep)
    jobs[1].poison = 1
    jobs[3].num_seeds = 300
    jobs[3].algorithm = 1  # flood with more than 254 labels
    res = vf.Batch().run(jobs)
    assert res[1]["status"] == -10 and res[3]["status"] == -4
    good = vf.Batch().run([jobs[i] for i in (0, 2, 4, 5)])
    assert [res[i]["checksum"] for i in (0, 2, 4, 5)] == [r["checksum"] for r in good]


def test_batch_marching_cubes_counts(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 3, 11, keep, edges=(48, 64), mc_every=1)
    res = vf.Batch().run(jobs)
    for m, r in zip(meta, res):
        ref, _ = oracle_job(m)
        mesh = O.marching_cubes(ref, *BOX)
        assert r["mc_triangles"] == mesh["tri"].shape[0] and r["mc_vertices"] == mesh["xyz"].shape[0]


def _production_reference(m, S0, loc, extra, extra_key, mdist, erode, frags=None, near=None):
    """CADScene::fractureModel on the oracle: fragment seeds (drawn, or the given impact seeds), nearSeeds
    when near = (impacts, seeds, spreading, key), mergeSeeds with the extra draw (the merged list follows
    the fragments), the fracture, then erosion or boundary detection."""
    g = O.occ_to_grid(O.voxelize(m["xyz"], m["tri"], *BOX, m["dims"]))
    if frags is None:
        frags = O.seed_uniform(g, S0, m["seed"], loc)
    if near:
        frags = O.seed_near(g, frags, *near)
    seeds = frags
    if extra:
        ex = O.seed_uniform(g, extra, extra_key, vf_SEED_BOTH)
        merged = O.seed_merge(frags, np.concatenate([frags, ex]), mdist)
        seeds = np.concatenate([frags, merged])
    if m["algo"] == 0:
        lab = O.naive(g, seeds, m["dist"])
    else:
        lab = O.flood(g, seeds, 6 if m["dist"] == 1 else 26, m["noise_seed"], m["amp"])
    if erode:
        kind, size, its, prob, thr, key = erode
        return O.erode(lab, kind, size, its, prob, thr, key)
    return O.detect_boundaries(lab, 1)


vf_SEED_BOTH = 2


def test_batch_and_server_production_seeding_and_erosion_match_oracle(vf):
    """The reference's default fracture sequence in batched and served jobs: fragment seeds on the surface,
    twice as many extra seeds merged as sub-seeds (the flood's sub-seed pass), erosion on some jobs."""
    keep = []
    jobs, meta = mixed_jobs(vf, 6, 51, keep, edges=(40, 72), locations=[vf.SEED_OUTER])
    erosions = [None, (1, 3, 3, 0.5, 0.5, 77), None, (0, 5, 2, 0.7, 0.4, 78), None, (2, 3, 1, 0.5, 0.5, 79)]
    cfg = []
    for i, (j, m, er) in enumerate(zip(jobs, meta, erosions)):
        j.extra_seeds = 2 * m["S"] if i != 4 else 0
        j.extra_seed = 1000 + i
        j.merge_distance = i % 3
        if er:
            j.erode, j.erosion_convolution, j.erosion_size, j.erosion_iterations = 1, er[0], er[1], er[2]
            j.erosion_probability, j.erosion_threshold, j.erosion_seed = er[3], er[4], er[5]
        cfg.append((m["S"], vf.SEED_OUTER, j.extra_seeds, j.extra_seed, j.merge_distance, er))
    with vf.Batch() as b:
        res = b.run(jobs)
    batched = [m["labels"].copy() for m in meta]
    for i, (m, r, c) in enumerate(zip(meta, res, cfg)):
        assert r["status"] == 0, (i, r)
        ref = _production_reference(m, *c)
        np.testing.assert_array_equal(batched[i], ref, err_msg=f"job {i}")
        np.testing.assert_array_equal(m["hist"], O.histogram(ref), err_msg=f"job {i}")
        assert r["checksum"] == O.checksum(ref)
    # the plain and served forms give the same grids
    plain = vf.plain_run(jobs)
    assert [p["checksum"] for p in plain] == [r["checksum"] for r in res]
    srv = vf.Server(64 << 30, max_slots=4)
    served = [srv.wait(t) for t in [srv.submit(j) for j in jobs]]
    srv.close()
    assert [x["checksum"] for x in served] == [r["checksum"] for r in res]
    for i, m in enumerate(meta):
        np.testing.assert_array_equal(m["labels"], batched[i])


def test_batch_and_server_impact_seeding_match_oracle(vf):
    """CADScene::fractureModel's impact paths in batched and served jobs: drawn fragment seeds followed by
    nearSeeds (numImpacts > 0), and the caller's impact seeds followed by one nearSeeds impact, some jobs
    with mergeSeeds after it.  nearSeeds' seed count is known only on the device."""
    keep = []
    jobs, meta = mixed_jobs(vf, 8, 61, keep, edges=(40, 72), locations=[vf.SEED_OUTER])
    cfg = []
    for i, (j, m) in enumerate(zip(jobs, meta)):
        frags = None
        if i % 2:  # the caller's seeds: surface cells of another draw, labels 2..S+1
            g = O.occ_to_grid(O.voxelize(m["xyz"], m["tri"], *BOX, m["dims"]))
            frags = np.ascontiguousarray(O.seed_uniform(g, m["S"], 500 + i, vf.SEED_OUTER))
            keep.append(frags)
            j.seeds = frags.ctypes.data
        j.impacts = 1 if i % 2 else 1 + i % 3
        j.impact_seeds = 4 + i
        j.impact_spreading = 2 + i % 3
        j.impact_seed = 900 + i
        if i in (2, 3, 6):
            j.extra_seeds = m["S"]
            j.extra_seed = 70 + i
            j.merge_distance = i % 3
        near = (j.impacts, j.impact_seeds, j.impact_spreading, j.impact_seed)
        cfg.append(((m["S"], vf.SEED_OUTER, j.extra_seeds, j.extra_seed, j.merge_distance, None), frags, near))
    refs = [_production_reference(m, *c, frags=f, near=nr) for m, (c, f, nr) in zip(meta, cfg)]
    with vf.Batch() as b:
        res = b.run(jobs)
        for i, (m, r) in enumerate(zip(meta, res)):
            print("impact job", i, m["dims"], m["algo"], m["S"], flush=True)
            assert r["status"] == 0, (i, r)
            np.testing.assert_array_equal(m["labels"], refs[i], err_msg=f"job {i}")
            np.testing.assert_array_equal(m["hist"], O.histogram(refs[i]), err_msg=f"job {i}")
        again = b.run(jobs)  # replay: nearSeeds runs again from the same fragments
    assert [x["checksum"] for x in again] == [r["checksum"] for r in res]
    plain = vf.plain_run(jobs)
    assert [p["checksum"] for p in plain] == [r["checksum"] for r in res]
    srv = vf.Server(64 << 30, max_slots=3)
    served = [srv.wait(t) for t in [srv.submit(j) for j in jobs]]
    srv.close()
    assert [x["status"] for x in served] == [0] * len(jobs)
    assert [x["checksum"] for x in served] == [r["checksum"] for r in res]
    # bad input is refused per job: a given seed outside the grid, spreading 0
    bad = np.zeros((meta[0]["S"], 4), np.uint32)
    bad[:, 0] = 4096
    bad[:, 3] = 2
    keep.append(bad)
    jobs[0].seeds = bad.ctypes.data
    jobs[2].impact_spreading = 0
    res = vf.Batch().run(jobs[:3])
    assert res[0]["status"] == -1 and res[2]["status"] == -1 and res[1]["status"] == 0


def _mesh_equal(m, r):
    for k in ("labels", "voff", "toff", "tri", "flag", "xyz"):
        np.testing.assert_array_equal(m[k], r[k], err_msg=k)


def test_batch_marching_cubes_meshes_match_oracle(vf):
    # every job's surfaces extracted in one launch set; meshes copied to the jobs' buffers; one job smoothed
    keep = []
    jobs, meta = mixed_jobs(vf, 4, 31, keep, edges=(40, 72), mc_every=1)
    it = 3
    sm = (it, 0.9, it, 0.2)
    bufs = []
    for i, j in enumerate(jobs):
        b = vf.MeshBuffers(400000, 800000)
        b.attach(j)
        bufs.append(b)
        if i == 2:
            j.smooth_iterations[:] = [sm[0], sm[2]]
            j.smooth_weights[:] = [sm[1], sm[3]]
    with vf.Batch() as bt:
        res = bt.run(jobs)
        res2 = bt.run(jobs)  # replay: the same meshes
    for i, (m, r, b) in enumerate(zip(meta, res, bufs)):
        assert r["status"] == 0, r
        ref, _ = oracle_job(m)
        want = O.marching_cubes(ref, *BOX, smooth=sm if i == 2 else None)
        got = b.mesh()
        assert got is not None, "mesh not copied"
        assert r["mc_triangles"] == want["tri"].shape[0] and r["mc_vertices"] == want["xyz"].shape[0]
        _mesh_equal(got, want)
    assert [r["mc_triangles"] for r in res2] == [r["mc_triangles"] for r in res]
    # a buffer too small: counts reported, nothing copied
    small = vf.MeshBuffers(10, 10)
    small.attach(jobs[0])
    r = vf.Batch().run([jobs[0]])[0]
    assert r["status"] == 0 and small.out.copied == 0 and small.out.num_triangles == res[0]["mc_triangles"]


def test_server_marching_cubes_equals_batch(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 6, 33, keep, edges=(40, 64), mc_every=1)
    bufs = [vf.MeshBuffers(300000, 600000) for _ in jobs]
    for b, j in zip(bufs, jobs):
        b.attach(j)
    srv = vf.Server(64 << 30, max_slots=3)
    out = [srv.wait(t) for t in [srv.submit(j) for j in jobs]]
    served = [b.mesh() for b in bufs]
    mem = srv.memory()
    srv.close()
    assert mem["max_unreserved_bytes"] == 0, mem
    for m, o, got in zip(meta, out, served):
        assert o["status"] == 0 and got is not None
        ref, _ = oracle_job(m)
        _mesh_equal(got, O.marching_cubes(ref, *BOX))


def test_server_results_match_alone_and_oracle(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 16, 5, keep)
    alone = vf.Batch().run(jobs)
    for m in meta:  # the served run below must write every output again
        m["labels"][...] = 0xDEAD
        m["hist"][...] = 0
    srv = vf.Server(64 << 30, max_slots=6)
    tickets = [srv.submit(j) for j in jobs]
    out = [srv.wait(t) for t in tickets]
    for a, o in zip(alone, out):
        assert o["status"] == 0 and o["checksum"] == a["checksum"] and o["occupied"] == a["occupied"]
        assert o["latency_ms"] >= o["queue_ms"] >= 0
    for m, o in zip(meta, out):  # served results against the oracle, cell for cell
        ref, occ = oracle_job(m)
        np.testing.assert_array_equal(m["labels"], ref)
        np.testing.assert_array_equal(m["hist"], O.histogram(ref))
        assert o["occupied"] == occ and o["checksum"] == O.checksum(ref)
    st = srv.stats()
    assert st["in_flight"] == 0 and st["reserved_bytes"] == 0 and st["completed"] == 16
    mem = srv.memory()
    assert mem["max_unreserved_bytes"] == 0, mem  # allocations never exceeded the admitted jobs' reservations
    srv.close()


def test_server_budget_covers_allocations_under_pressure(vf):
    # a budget that admits only a few of these jobs at once: the scheduler's allocations (job blocks,
    # tables, worklists, voxelizer and seed scratch) must stay within what the admitted jobs reserved
    keep = []
    jobs, meta = mixed_jobs(vf, 12, 17, keep, edges=(64, 128))
    alone = vf.Batch().run(jobs)
    srv = vf.Server(160 << 20, max_slots=8)
    tickets = [srv.submit(j) for j in jobs]
    out = [srv.wait(t) for t in tickets]
    assert [o["checksum"] for o in out] == [a["checksum"] for a in alone]
    mem = srv.memory()
    assert mem["max_unreserved_bytes"] == 0, mem
    assert mem["reserved_bytes"] == 0
    assert mem["peak_device_bytes"] <= 160 << 20, mem  # cached blocks included
    srv.close()


def test_batch_rereads_host_meshes_refilled_in_place(vf):
    # same descriptors, new mesh contents: the batch uploads the host meshes again on every run
    keep = []
    jobs, meta = mixed_jobs(vf, 3, 41, keep, edges=(48, 64))
    for j, m in zip(jobs, meta):
        assert j.xyz == m["xyz"].ctypes.data  # the job points at the test's own buffer
    b = vf.Batch()
    first = b.run(jobs)
    for m in meta:
        m["xyz"][:, 0] *= np.float32(0.6)  # squeeze the vessels in x, in place
    second = b.run(jobs)
    for m, r1, r2 in zip(meta, first, second):
        assert r2["status"] == 0
        ref, occ = oracle_job(m)
        np.testing.assert_array_equal(m["labels"], ref)
        assert r2["occupied"] == occ and r1["occupied"] != occ
    b.close()


def test_device_mesh_with_bad_index_fails_alone(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 3, 43, keep, edges=(40, 56))
    tri = meta[1]["tri"].copy()
    tri[5, 1] = meta[1]["xyz"].shape[0]  # one past the last vertex
    dev = vf.DeviceMesh(meta[1]["xyz"], tri)
    bad = dev.job(meta[1]["dims"], *BOX, meta[1]["S"], meta[1]["seed"], meta[1]["algo"], meta[1]["dist"],
                  noise_seed=meta[1]["noise_seed"], noise_amp=meta[1]["amp"], boundary_size=1)
    res = vf.Batch().run([jobs[0], bad, jobs[2]])
    assert res[1]["status"] == -1
    good = vf.Batch().run([jobs[0], jobs[2]])
    assert [res[0]["checksum"], res[2]["checksum"]] == [r["checksum"] for r in good]
    with vf.Grid(meta[1]["dims"], *BOX) as g:
        with pytest.raises(vf.VfError):
            g.voxelize_device(dev)
        with pytest.raises(vf.VfError):
            g.voxelize(meta[1]["xyz"], tri)
    dev.close()


def test_server_budget_poison_cancel(vf):
    keep = []
    jobs, meta = mixed_jobs(vf, 6, 9, keep)
    big = vf.make_job(meta[0]["xyz"], meta[0]["tri"], (1024, 1024, 1024), *BOX, 4, 1, 1, 2, keep=keep)
    jobs[2].poison = 1
    srv = vf.Server(2 << 30, max_slots=2)
    t_big = srv.submit(big)
    tickets = [srv.submit(j) for j in jobs]
    assert srv.wait(t_big)["status"] == -5
    out = [srv.wait(t) for t in tickets]
    assert out[2]["status"] == -10
    ok = [o for i, o in enumerate(out) if i != 2]
    assert all(o["status"] == 0 for o in ok)
    ref = vf.Batch().run([jobs[i] for i in range(6) if i != 2])
    assert [o["checksum"] for o in ok] == [r["checksum"] for r in ref]
    st = srv.stats()
    assert st["in_flight"] == 0 and st["reserved_bytes"] == 0
    srv.close()
