"""Parity at the headline configurations' full sizes.

128^3 naive (8 seeds, Manhattan) and 256^3 flood (32 seeds, Euclidean ->
26 neighbours, noise amplitude 2, histogram) are compared with the oracle
cell for cell.  At 512^3 (~1M-triangle vessel, 128 seeds, Chebyshev) naive,
flood (noise 2), boundaries, histogram and every fragment's marching-cubes
mesh are compared with the oracle (FP32 vertices exactly, and within 1e-5 of
the FP64 twin), the fragments are checked to be connected and the meshes
closed and consistently oriented, and the batched form must equal the
single-grid path.  The full 64-job mixed batch is compared with the oracle
job by job.  Grids beyond 2^30 cells (flood keys beyond 4 GiB) and grids
whose Euclidean float roots collide are checked against the oracle on
translated and sparse inputs."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BOX = ([-1.03, -1.03, -1.03], [1.03, 1.03, 1.03])
EUC, MAN, CHE = 0, 1, 2


def vessel(vf, seed, triangles):
    return vf.vessel_mesh(seed, *vf.vessel_segments(triangles))


@pytest.mark.timeout(600)
def test_cfg_128_naive_manhattan_exact(vf):
    xyz, tri = vessel(vf, 1, 20000)
    with vf.Grid((128, 128, 128), *BOX) as g:
        g.voxelize(xyz, tri)
        base = g.read()
        seeds = g.seed_uniform(8, 1)
        g.fracture_naive(seeds, MAN)
        got = g.read()
    ref_base = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, (128, 128, 128)))
    np.testing.assert_array_equal(base, ref_base)
    np.testing.assert_array_equal(seeds, O.seed_uniform(ref_base, 8, 1))
    np.testing.assert_array_equal(got, O.naive(ref_base, seeds, MAN))


@pytest.mark.timeout(600)
def test_cfg_256_flood_noise_histogram_exact(vf):
    xyz, tri = vessel(vf, 2, 200000)
    with vf.Grid((256, 256, 256), *BOX) as g:
        g.voxelize(xyz, tri)
        base = g.read()
        seeds = g.seed_uniform(32, 2)
        g.fracture_flood(seeds, EUC, noise_seed=2, noise_amp=2)
        got = g.read()
        hist = g.histogram()
        meta = g.fragment_metadata()
    ref_base = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, (256, 256, 256)))
    np.testing.assert_array_equal(base, ref_base)
    ref = O.flood(ref_base, seeds, 26, 2, 2)
    np.testing.assert_array_equal(got, ref)
    ref_hist = O.histogram(ref)
    np.testing.assert_array_equal(hist, ref_hist)
    keep = np.flatnonzero(ref_hist[2:]) + 2
    total = int(ref_hist[2:].sum())
    assert [m["value"] for m in meta] == keep.tolist()
    assert [m["voxels"] for m in meta] == ref_hist[keep].tolist()
    assert all(m["occupied_voxels"] == total for m in meta)
    assert [m["percentage"] for m in meta] == (ref_hist[keep].astype(np.float32) / np.float32(total)).tolist()


def fragments_connected(lab, labels):
    from scipy import ndimage
    objs = ndimage.find_objects(lab.astype(np.int32))
    for L in labels:
        sl = objs[L - 1]
        assert sl is not None, f"label {L} missing"
        _, n = ndimage.label(lab[sl] == L, structure=np.ones((3, 3, 3), bool))
        assert n == 1, f"label {L} has {n} components"


def closed_and_oriented(m):
    for i in range(len(m["labels"])):
        t = m["tri"][m["toff"][i]:m["toff"][i + 1]].astype(np.uint64)
        d = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
        key = d[:, 0] * np.uint64(1 << 32) + d[:, 1]
        assert np.unique(key).size == key.size, "a directed edge is used twice"
        rev = d[:, 1] * np.uint64(1 << 32) + d[:, 0]
        assert np.all(np.isin(rev, key)), "an edge has one triangle only"


def compare_mesh(m, r):
    np.testing.assert_array_equal(m["labels"], r["labels"])
    np.testing.assert_array_equal(m["voff"], r["voff"])
    np.testing.assert_array_equal(m["toff"], r["toff"])
    np.testing.assert_array_equal(m["tri"], r["tri"])
    np.testing.assert_array_equal(m["flag"], r["flag"])
    np.testing.assert_array_equal(m["xyz"], r["xyz"])
    scale = np.maximum(np.abs(r["xyz64"]), 2.06)
    assert np.all(np.abs(m["xyz"].astype(np.float64) - r["xyz64"]) <= 1e-5 * scale)


@pytest.mark.timeout(900)
def test_cfg_512_chebyshev_naive_flood_mc_exact(vf):
    from scipy import ndimage
    dims = (512, 512, 512)
    xyz, tri = vessel(vf, 3, 1000000)
    with vf.Grid(dims, *BOX) as g:
        g.voxelize(xyz, tri)
        base = g.read()
        seeds = g.seed_uniform(128, 3)
        g.fracture_naive(seeds, CHE)
        naive = g.read()
        g.write(base)
        g.fracture_flood(seeds, CHE, noise_seed=3, noise_amp=2)
        flood = g.read()
        g.detect_boundaries(1)
        bnd = g.read()
        hist = g.histogram()
        m = g.marching_cubes()
    print("512^3: occupied", int((base != 0).sum()), "triangles", m["tri"].shape[0], flush=True)
    ref_base = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, dims))
    np.testing.assert_array_equal(base, ref_base)
    np.testing.assert_array_equal(seeds, O.seed_uniform(ref_base, 128, 3))
    np.testing.assert_array_equal(naive, O.naive(ref_base, seeds, CHE))
    del naive
    ref_flood = O.flood(ref_base, seeds, 26, 3, 2)
    np.testing.assert_array_equal(flood, ref_flood)
    ref_bnd = O.detect_boundaries(ref_flood, 1)
    np.testing.assert_array_equal(bnd, ref_bnd)
    np.testing.assert_array_equal(hist, O.histogram(ref_bnd))
    r = O.marching_cubes(ref_bnd, *BOX)
    assert len(r["labels"]) == 128
    compare_mesh(m, r)
    del r

    labels = seeds[:, 3].astype(np.int64) & 0xFF
    labelled = flood > 1
    free = flood == 1
    near = ndimage.binary_dilation(labelled, structure=np.ones((3, 3, 3), bool))
    assert not np.any(near & free), "a FREE cell touches a fragment"
    fragments_connected(flood, labels.tolist())
    closed_and_oriented(m)

    # the batched form equals the single-grid path
    dev = vf.DeviceMesh(xyz, tri)
    out = np.empty(dims, np.uint16)
    job = dev.job(dims, *BOX, 128, 3, 1, CHE, noise_seed=3, noise_amp=2, boundary_size=0)
    job.labels_out = out.ctypes.data
    with vf.Batch() as b:
        res = b.run([job])[0]
    print("batched 512^3 flood:", res, flush=True)
    assert res["status"] == 0
    assert res["occupied"] == int((base != 0).sum())
    assert res["num_labels"] == 128
    np.testing.assert_array_equal(out, flood)
    dev.close()


def _oracle_bench_job(spec, xyz, tri):
    e = spec["edge"]
    occ = O.voxelize(xyz, tri, *BOX, (e, e, e))
    g = O.occ_to_grid(occ)
    seeds = O.seed_uniform(g, spec["seeds"], spec["seed_key"])
    if spec["algo"] == 0:
        lab = O.naive(g, seeds, spec["dist"])
    else:
        lab = O.flood(g, seeds, 6 if spec["dist"] == 1 else 26, spec["noise_seed"], 2)
    return O.detect_boundaries(lab, 1), int(occ.sum())


@pytest.mark.timeout(900)
def test_cfg4_mixed_batch_64_jobs_match_oracle(vf):
    """The benchmark's mixed batch (64 jobs, edges U[64, 384], 5k-300k triangles, 2-64 seeds, distance
    uniform, naive/flood 50:50, flood noise 2, boundaries + histogram), every job against the oracle."""
    import concurrent.futures as cf
    import bench
    specs = bench.job_specs(64, 4)
    meshes, jobs = bench.build_jobs(vf, specs)
    outs = []
    for s, j in zip(specs, jobs):
        lab = np.empty((s["edge"],) * 3, np.uint16)
        hist = np.empty(65536, np.uint32)
        j.labels_out = lab.ctypes.data
        j.histogram_out = hist.ctypes.data
        outs.append((lab, hist))
    with vf.Batch() as b:
        res = b.run(jobs)
    host = [vf.vessel_mesh(s["mesh_seed"], *vf.vessel_segments(s["tris"])) for s in specs]
    print("cfg4: 64 jobs run on the GPU; oracle on", min(16, len(specs)), "threads", flush=True)
    with cf.ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(lambda a: _oracle_bench_job(*a), [(s, *h) for s, h in zip(specs, host)]))
    for i, (s, r, (lab, hist), (ref, occ)) in enumerate(zip(specs, res, outs, refs)):
        assert r["status"] == 0, (i, r)
        np.testing.assert_array_equal(lab, ref, err_msg=f"job {i} {s}")
        np.testing.assert_array_equal(hist, O.histogram(ref), err_msg=f"job {i}")
        assert r["occupied"] == occ and r["checksum"] == O.checksum(ref), i
    for m in meshes:
        m.close()


@pytest.mark.timeout(600)
def test_flood_and_naive_beyond_2_30_cells_match_oracle(vf):
    """Keys of a grid over 2^30 cells span more than 4 GiB: a blob placed at the grid's far corner is
    fractured as the same blob alone (no noise, so the result does not depend on where it sits)."""
    rng = np.random.default_rng(30)
    small = (48, 40, 36)
    occ = (rng.random(small) < 0.62).astype(np.uint8)
    g0 = O.occ_to_grid(occ)
    seeds = O.seed_uniform(g0, 6, 31)
    ref_flood = O.flood(g0, seeds, 26, 0, 0)
    ref_naive = O.naive(g0, seeds, EUC)
    dims = (1040, 1040, 1040)
    assert dims[0] * dims[1] * dims[2] > (1 << 30)
    off = np.array([990, 1000, 1002])
    big = np.zeros(dims, np.uint16)
    sl = tuple(slice(o, o + n) for o, n in zip(off, small))
    big[sl] = g0
    bseeds = seeds.copy()
    bseeds[:, :3] += off.astype(np.uint32)
    with vf.Grid(dims, *BOX) as g:
        g.write(big)
        g.fracture_flood(bseeds, CHE, noise_seed=0, noise_amp=0)
        got = g.read()
        np.testing.assert_array_equal(got[sl], ref_flood)
        got[sl] = 0
        assert not got.any()
        g.write(big)
        g.fracture_naive(bseeds, EUC)
        got = g.read()
        np.testing.assert_array_equal(got[sl], ref_naive)


def test_naive_euclid_float_roots_match_oracle(vf):
    """Squared distances beyond 4197200 can share a float root: the reference's float sqrt ties them,
    and the first seed wins."""
    grid = np.zeros((2048, 2048, 1), np.uint16)
    grid[0, 0, 0] = 1
    two = np.array([[659, 1951, 0, 2], [616, 1965, 0, 3]], np.uint32)
    with vf.Grid(grid.shape, *BOX) as g:
        g.write(grid)
        g.fracture_naive(two, EUC)
        assert g.read()[0, 0, 0] == 2
    rng = np.random.default_rng(11)
    shape = (2048, 2048, 2)
    grid = np.zeros(shape, np.uint16)
    cells = rng.integers(0, [2048, 2048, 2], size=(4000, 3))
    grid[cells[:, 0], cells[:, 1], cells[:, 2]] = 1
    occ = np.argwhere(grid != 0)
    seeds = occ[rng.choice(len(occ), 96, replace=False)]
    seeds = seeds[np.lexsort((seeds[:, 2], seeds[:, 1], seeds[:, 0]))]
    seeds = np.concatenate([seeds, (2 + np.arange(96))[:, None]], axis=1).astype(np.uint32)
    for dist in (EUC, MAN, CHE):
        with vf.Grid(shape, *BOX) as g:
            g.write(grid)
            g.fracture_naive(seeds, dist)
            np.testing.assert_array_equal(g.read(), O.naive(grid, seeds, dist))


def test_batched_naive_euclid_float_roots_match_oracle(vf):
    # the batched engine picks the same key form: a grid whose squared distances pass 4197200
    dims = (2048, 1536, 8)
    xyz, tri = vessel(vf, 7, 20000)
    lab = np.empty(dims, np.uint16)
    dev = vf.DeviceMesh(xyz, tri)
    job = dev.job(dims, *BOX, 24, 7, 0, EUC, boundary_size=0)
    job.labels_out = lab.ctypes.data
    with vf.Batch() as b:
        r = b.run([job])[0]
    assert r["status"] == 0
    g = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, dims))
    seeds = O.seed_uniform(g, 24, 7)
    np.testing.assert_array_equal(lab, O.naive(g, seeds, EUC))
    dev.close()
