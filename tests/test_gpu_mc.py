"""GPU marching cubes against the oracle: labels, offsets, welded vertices and
triangles must match exactly; FP32 vertices within 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BOX = ([-1.03, -1.03, -1.03], [1.03, 1.03, 1.03])


def compare(m, r):
    np.testing.assert_array_equal(m["labels"], r["labels"])
    np.testing.assert_array_equal(m["voff"], r["voff"])
    np.testing.assert_array_equal(m["toff"], r["toff"])
    np.testing.assert_array_equal(m["tri"], r["tri"])
    np.testing.assert_array_equal(m["flag"], r["flag"])
    np.testing.assert_array_equal(m["xyz"], r["xyz"])
    scale = np.maximum(np.abs(r["xyz64"]), 2.06)
    assert np.all(np.abs(m["xyz"].astype(np.float64) - r["xyz64"]) <= 1e-5 * scale)


@pytest.mark.parametrize("dims,S", [((64, 64, 64), 12), ((45, 52, 37), 5)])
def test_mc_fractured_vessel(vf, dims, S):
    xyz, tri = vf.vessel_mesh(2, *vf.vessel_segments(12000))
    g = vf.Grid(dims, *BOX)
    g.voxelize(xyz, tri)
    base = g.read()
    seeds = O.seed_uniform(base, S, 3)
    g.fracture_naive(seeds, 0)
    g.detect_boundaries(1)
    grid = g.read()
    m = g.marching_cubes()
    r = O.marching_cubes(grid, *BOX)
    assert len(r["labels"]) == S
    compare(m, r)
    # single-label extraction equals that label's slice of the all-label mesh
    L = int(r["labels"][1])
    m1 = g.marching_cubes(L)
    r1 = O.marching_cubes(grid, *BOX, label=L)
    compare(m1, r1)


def test_mc_random_labels_and_box(vf):
    rng = np.random.default_rng(6)
    dims = (20, 18, 16)
    grid = np.zeros(dims, np.uint16)
    grid[2:18, 2:16, 2:14] = rng.integers(2, 9, size=(16, 14, 12))
    grid[rng.random(dims) < 0.05] |= 0x8000
    g = vf.Grid(dims, [-1.3, -0.7, 0.2], [2.1, 1.9, 3.3])
    g.write(grid)
    m = g.marching_cubes()
    compare_box = O.marching_cubes(grid, [-1.3, -0.7, 0.2], [2.1, 1.9, 3.3])
    np.testing.assert_array_equal(m["tri"], compare_box["tri"])
    np.testing.assert_array_equal(m["xyz"], compare_box["xyz"])
    np.testing.assert_array_equal(m["flag"], compare_box["flag"])
    box = np.zeros((9, 8, 10), np.uint16)
    box[2:6, 2:5, 2:9] = 7
    g2 = vf.Grid(box.shape, [0, 0, 0], list(box.shape))
    g2.write(box)
    m2 = g2.marching_cubes()
    a, b, c = 4, 3, 7
    T = 4 * ((a - 1) * (b - 1) + (b - 1) * (c - 1) + (c - 1) * (a - 1)) + 8 * ((a - 1) + (b - 1) + (c - 1)) + 8
    assert m2["tri"].shape[0] == T and m2["xyz"].shape[0] == 2 + T // 2


def test_mc_empty_grid(vf):
    g = vf.Grid((10, 10, 10), *BOX)
    m = g.marching_cubes()
    assert m["tri"].shape[0] == 0 and m["labels"].shape[0] == 0


@pytest.mark.parametrize("dims,S,dist", [((64, 64, 64), 10, 0), ((50, 44, 58), 6, 1)])
def test_mc_smoothing_matches_oracle(vf, dims, S, dist):
    xyz, tri = vf.vessel_mesh(5, *vf.vessel_segments(12000))
    g = vf.Grid(dims, *BOX)
    g.voxelize(xyz, tri)
    seeds = O.seed_uniform(g.read(), S, 7)
    g.fracture_naive(seeds, dist)
    g.detect_boundaries(1)
    grid = g.read()
    it = int(np.float32(max(dims)) * np.float32(0.048))  # the reference's default iteration count
    sm = (it, 0.9, it, 0.2)
    g.marching_cubes()
    g.smooth_mesh(*sm)
    m = g.read_mesh()
    r = O.marching_cubes(grid, *BOX, smooth=sm)
    np.testing.assert_array_equal(m["tri"], r["tri"])
    np.testing.assert_array_equal(m["xyz"], r["xyz"])
    g.close()


def decode_binm(data):
    """The layout CADModel::saveBinary writes (CADModel.cpp:839-855): u32 vertex count, 64-byte
    vertices (position first), u32 face count, 16-byte faces (3 indices, component id)."""
    nv = int(np.frombuffer(data, np.uint32, 1, 0)[0])
    vert = np.frombuffer(data, np.float32, 16 * nv, 4).reshape(nv, 16)
    o = 4 + 64 * nv
    nt = int(np.frombuffer(data, np.uint32, 1, o)[0])
    face = np.frombuffer(data, np.uint32, 4 * nt, o + 4).reshape(nt, 4)
    assert len(data) == o + 4 + 16 * nt
    return vert, face


def test_mc_binm_export(vf):
    xyz, tri = vf.vessel_mesh(3, *vf.vessel_segments(12000))
    dims = (48, 40, 56)
    g = vf.Grid(dims, *BOX)
    g.voxelize(xyz, tri)
    seeds = O.seed_uniform(g.read(), 5, 9)
    g.fracture_flood(seeds, 0)
    g.detect_boundaries(1)
    grid = g.read()
    it = int(np.float32(max(dims)) * np.float32(0.048))
    sm = (it, 0.9, it, 0.2)
    g.marching_cubes()
    g.smooth_mesh(*sm)
    r = O.marching_cubes(grid, *BOX, smooth=sm)
    assert len(r["labels"]) == 5
    for i in range(len(r["labels"])):
        vert, face = decode_binm(g.export_binm(i))
        v0, v1, t0, t1 = r["voff"][i], r["voff"][i + 1], r["toff"][i], r["toff"][i + 1]
        np.testing.assert_array_equal(vert[:, :3], r["xyz"][v0:v1])
        assert not vert[:, 3:].any()  # dataset export: normal, texcoord, tangent and padding stay zero
        np.testing.assert_array_equal(face[:, :3], r["tri"][t0:t1])
        assert not face[:, 3].any()
    with pytest.raises(vf.VfError):
        g.export_binm(5)
    g.close()
