"""Parity at the headline configurations' full sizes.

128^3 naive (8 seeds, Manhattan) and 256^3 flood (32 seeds, Euclidean ->
26 neighbours, noise amplitude 2, histogram) are compared with the oracle
cell for cell.  At 512^3 (~1M-triangle vessel, 128 seeds, Chebyshev) naive
is compared cell for cell; the flood, whose oracle run is too slow for a
test, is checked through properties that do not depend on size: labels are
the seeds' labels, every seed keeps its own, EMPTY cells are exactly the
unoccupied ones, no FREE cell touches a labelled cell (the growth is
closed), every fragment is one 26-connected component, and the batched form
equals the single-grid path; its marching-cubes meshes are closed and
consistently oriented."""
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
    ref_base = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, (256, 256, 256)))
    np.testing.assert_array_equal(base, ref_base)
    ref = O.flood(ref_base, seeds, 26, 2, 2)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(hist, O.histogram(ref))


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


@pytest.mark.timeout(900)
def test_cfg_512_chebyshev_naive_exact_flood_properties(vf):
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
        m = g.marching_cubes()
    print("512^3: occupied", int((base != 0).sum()), flush=True)
    ref_base = O.occ_to_grid(O.voxelize(xyz, tri, *BOX, dims))
    np.testing.assert_array_equal(base, ref_base)
    np.testing.assert_array_equal(naive, O.naive(ref_base, seeds, CHE))
    del naive

    labels = seeds[:, 3].astype(np.int64) & 0xFF
    occupied = base != 0
    assert np.array_equal(flood != 0, occupied)
    assert set(np.unique(flood).tolist()) <= set(labels.tolist()) | {0, 1}
    assert np.all(flood[seeds[:, 0], seeds[:, 1], seeds[:, 2]] == labels)
    labelled = flood > 1
    free = flood == 1
    near = ndimage.binary_dilation(labelled, structure=np.ones((3, 3, 3), bool))
    assert not np.any(near & free), "a FREE cell touches a fragment"
    fragments_connected(flood, labels.tolist())
    closed_and_oriented(m)
    assert len(m["labels"]) == len(labels)

    # the batched form equals the single-grid path
    dev = vf.DeviceMesh(xyz, tri)
    out = np.empty(dims, np.uint16)
    job = dev.job(dims, *BOX, 128, 3, 1, CHE, noise_seed=3, noise_amp=2, boundary_size=0)
    job.labels_out = out.ctypes.data
    with vf.Batch() as b:
        r = b.run([job])[0]
    assert r["status"] == 0
    np.testing.assert_array_equal(out, flood)
    dev.close()
