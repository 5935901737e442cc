"""GPU parity: every stage of libvfrag, called through the C ABI, against the
CPU oracle on the same seeded inputs.  Integer outputs must be bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EUC, MAN, CHE = 0, 1, 2
BOX = ([-1.03, -1.03, -1.03], [1.03, 1.03, 1.03])


def vessel(vf, seed, tris):
    r, h = vf.vessel_segments(tris)
    return vf.vessel_mesh(seed, r, h)


def voxelized(vf, dims, seed=1, tris=20000):
    xyz, tri = vessel(vf, seed, tris)
    g = vf.Grid(dims, *BOX)
    g.voxelize(xyz, tri)
    return g, xyz, tri


@pytest.mark.parametrize("dims", [(64, 64, 64), (96, 80, 72), (37, 41, 29), (128, 128, 128)])
def test_voxelize_matches_oracle(vf, dims):
    print("voxelize", dims, flush=True)
    g, xyz, tri = voxelized(vf, dims, seed=sum(dims), tris=8000)
    occ = g.occupancy()
    ref = O.voxelize(xyz, tri, *BOX, dims)
    assert ref.sum() > 0
    np.testing.assert_array_equal(occ, ref)
    np.testing.assert_array_equal(g.read(), O.occ_to_grid(ref))
    assert g.count_occupied() == int(ref.sum())


def test_set_aabb_then_voxelize_matches_oracle(vf):
    # RegularGrid::setAABB: a new box and new dims on an existing grid (cells cleared, cell size = box / dims)
    xyz, tri = vessel(vf, 12, 6000)
    g = vf.Grid((40, 40, 40), *BOX)
    g.voxelize(xyz, tri)
    mn, mx, dims = [-1.2, -0.9, -1.1], [0.8, 1.15, 0.95], (52, 36, 60)
    g.set_aabb(mn, mx, dims)
    assert g.dims == dims
    assert not g.read().any()  # cleared
    g.voxelize(xyz, tri)
    ref = O.voxelize(xyz, tri, mn, mx, dims)
    assert ref.sum() > 0
    np.testing.assert_array_equal(g.occupancy(), ref)
    seeds = g.seed_uniform(5, 13)
    g.fracture_naive(seeds, 0)
    np.testing.assert_array_equal(g.read(), O.naive(O.occ_to_grid(ref), seeds, 0))
    g.close()


def test_voxelize_merges_into_nonempty_grid(vf):
    dims = (48, 48, 48)
    g, xyz, tri = voxelized(vf, dims, seed=3, tris=4000)
    base = np.zeros(dims, np.uint16)
    base[:4] = 7
    g.write(base)
    g.voxelize(xyz, tri)
    ref = O.voxelize(xyz, tri, *BOX, dims)
    expect = base.copy()
    expect[ref > 0] = 1
    np.testing.assert_array_equal(g.read(), expect)


@pytest.mark.parametrize("loc", [0, 1, 2])
def test_seed_uniform_matches_oracle(vf, loc):
    g, _, _ = voxelized(vf, (64, 64, 64), seed=5)
    ref_grid = g.read()
    for n, key in [(8, 1), (32, 2), (128, 3)]:
        try:
            ref = O.seed_uniform(ref_grid, n, key, loc)
        except ValueError:  # fewer candidates than seeds
            with pytest.raises(vf.VfError) as e:
                g.seed_uniform(n, key, loc)
            assert e.value.code == -9
            continue
        np.testing.assert_array_equal(g.seed_uniform(n, key, loc), ref)


@pytest.mark.parametrize("dist", [EUC, MAN, CHE])
@pytest.mark.parametrize("dims", [(64, 64, 64), (50, 45, 43)])
def test_naive_matches_oracle(vf, dist, dims):
    print("naive", dist, dims, flush=True)
    g, _, _ = voxelized(vf, dims, seed=7)
    base = g.read()
    for S in (1, 8, 33):
        seeds = O.seed_uniform(base, S, 11 + S)
        g.write(base)
        g.fracture_naive(seeds, dist)
        np.testing.assert_array_equal(g.read(), O.naive(base, seeds, dist))


def test_naive_many_seeds_and_64bit_keys(vf):
    # 2000 seeds on a 1500x1500 slab: squared-Euclidean keys need 64 bits
    dims = (1500, 1500, 2)
    rng = np.random.default_rng(3)
    base = (rng.random(dims) < 0.02).astype(np.uint16)
    g = vf.Grid(dims, *BOX)
    g.write(base)
    seeds = O.seed_uniform(base, 2000, 9)
    for dist in (EUC, MAN, CHE):
        g.write(base)
        g.fracture_naive(seeds, dist)
        np.testing.assert_array_equal(g.read(), O.naive(base, seeds, dist))


@pytest.mark.parametrize("dist", [EUC, MAN, CHE])
def test_naive_dense_tiles(vf, dist):
    # every cell occupied: 2048 cells per tile, 8 per lane, beside the 1-4 per lane of sparse grids
    print("naive dense", dist, flush=True)
    dims = (40, 24, 36)
    base = np.ones(dims, np.uint16)
    g = vf.Grid(dims, *BOX)
    for S in (9, 37):
        seeds = O.seed_uniform(base, S, 5 + S)
        g.write(base)
        g.fracture_naive(seeds, dist)
        np.testing.assert_array_equal(g.read(), O.naive(base, seeds, dist))


def test_naive_ties_and_seeds_outside_rejected(vf):
    g = vf.Grid((5, 1, 8), *BOX)
    g.write(np.ones((5, 1, 8), np.uint16))
    seeds = np.array([[4, 0, 0, 7], [0, 0, 0, 9]], np.uint32)
    g.fracture_naive(seeds, MAN)
    np.testing.assert_array_equal(g.read(), O.naive(np.ones((5, 1, 8), np.uint16), seeds, MAN))
    with pytest.raises(vf.VfError):
        g.fracture_naive(np.array([[5, 0, 0, 2]], np.uint32), MAN)


@pytest.mark.parametrize("nb,amp", [(6, 0), (26, 0), (6, 2), (26, 2), (26, 1)])
@pytest.mark.parametrize("dims", [(64, 64, 64), (61, 47, 53)])
def test_flood_matches_oracle(vf, nb, amp, dims):
    print("flood", nb, amp, dims, flush=True)
    g, _, _ = voxelized(vf, dims, seed=9)
    base = g.read()
    seeds = O.seed_uniform(base, 16, 21)
    g.fracture_flood(seeds, CHE, nb, noise_seed=1234, noise_amp=amp)
    np.testing.assert_array_equal(g.read(), O.flood(base, seeds, nb, 1234, amp))


def test_flood_distance_selects_neighbourhood(vf):
    g, _, _ = voxelized(vf, (48, 48, 48), seed=4)
    base = g.read()
    seeds = O.seed_uniform(base, 6, 5)
    g.fracture_flood(seeds, MAN)  # FloodFracturer.cpp:114: Manhattan -> 6 neighbours
    np.testing.assert_array_equal(g.read(), O.flood(base, seeds, 6))
    g.write(base)
    g.fracture_flood(seeds, EUC)  # anything else -> 26
    np.testing.assert_array_equal(g.read(), O.flood(base, seeds, 26))


def test_flood_edge_cases(vf):
    dims = (20, 20, 20)
    rng = np.random.default_rng(1)
    base = (rng.random(dims) < 0.6).astype(np.uint16)
    base[10] = 0  # unreachable half
    g = vf.Grid(dims, *BOX)
    g.write(base)
    seeds = np.array([[2, 2, 2, 2], [2, 2, 2, 3], [3, 3, 3, 4], [15, 15, 15, 5]], np.uint32)
    base_h = base.copy()
    base_h[15, 15, 15] = 0  # a seed on an EMPTY cell is still a source
    g.write(base_h)
    g.fracture_flood(seeds, CHE, 26, noise_seed=5, noise_amp=2)
    np.testing.assert_array_equal(g.read(), O.flood(base_h, seeds, 26, 5, 2))
    for bad in (0x101, 0x10002):  # fragment id 1 (FREE); beyond 16 bits
        with pytest.raises(vf.VfError) as e:
            g.fracture_flood(np.array([[0, 0, 0, bad]], np.uint32), CHE)
        assert e.value.code == -4
    # no seeds: every occupied cell stays FREE
    g.write(base)
    g.fracture_flood(np.zeros((0, 4), np.uint32), CHE)
    np.testing.assert_array_equal(g.read(), O.flood(base, np.zeros((0, 4), np.uint32), 26))


@pytest.mark.parametrize("nb", [6, 26])
def test_flood_seed_in_empty_brick(vf, nb):
    """Seeds on EMPTY cells of 8^3 bricks without an occupied cell: the key init writes no keys there,
    so the seed kernel gives such a brick its keys; one seed touches the occupied block across the brick
    boundary at x = 16 and grows into it, one sits alone."""
    dims = (40, 36, 44)
    base = np.zeros(dims, np.uint16)
    base[:16, 4:30, 2:40] = 1
    seeds = np.array([[5, 10, 10, 2], [16, 12, 20, 3], [33, 30, 41, 4], [0, 4, 2, 5]], np.uint32)
    g = vf.Grid(dims, *BOX)
    g.write(base)
    g.fracture_flood(seeds, MAN if nb == 6 else CHE, nb, noise_seed=9, noise_amp=1)
    out = g.read()
    ref = O.flood(base, seeds, nb, 9, 1)
    np.testing.assert_array_equal(out, ref)
    assert out[33, 30, 41] == 4 and (out == 3).sum() > 100
    g.close()


@pytest.mark.parametrize("size", [1, 2])
def test_boundaries_histogram_undo(vf, size):
    g, _, _ = voxelized(vf, (64, 64, 64), seed=13)
    base = g.read()
    seeds = O.seed_uniform(base, 24, 4)
    g.fracture_naive(seeds, EUC)
    lab = O.naive(base, seeds, EUC)
    g.detect_boundaries(size)
    bnd = O.detect_boundaries(lab, size)
    np.testing.assert_array_equal(g.read(), bnd)
    np.testing.assert_array_equal(g.histogram(), O.histogram(bnd))
    g.undo_mask()
    np.testing.assert_array_equal(g.read(), O.undo_mask(bnd))
    g.homogenize()
    np.testing.assert_array_equal(g.read(), O.homogenize(lab))


def test_histogram_wide_labels(vf):
    rng = np.random.default_rng(8)
    dims = (32, 32, 40)
    grid = rng.integers(0, 65536, size=dims, dtype=np.uint16)
    g = vf.Grid(dims, *BOX)
    g.write(grid)
    np.testing.assert_array_equal(g.histogram(), O.histogram(grid))
    np.testing.assert_array_equal(g.histogram(300), O.histogram(grid)[:300])


def test_remove_isolated_matches_oracle(vf):
    g, _, _ = voxelized(vf, (64, 64, 64), seed=17)
    base = g.read()
    seeds = O.seed_uniform(base, 12, 3)
    g.fracture_naive(seeds, EUC)
    lab = O.naive(base, seeds, EUC)
    g.remove_isolated_regions(seeds)
    np.testing.assert_array_equal(g.read(), O.remove_isolated(lab, seeds))


def test_rle_fixture_roundtrip(vf):
    from pathlib import Path
    data = (Path(__file__).parent / "golden" / "AL_12B_grid_128r.rle").read_bytes()
    ref = O.rle_decode(data)
    g = vf.Grid(ref.shape, *BOX)
    g.import_rle(data)
    np.testing.assert_array_equal(g.read(), ref)
    assert g.export_rle() == data


def test_tiny_and_empty_grids(vf):
    g = vf.Grid((1, 1, 1), *BOX)
    assert g.count_occupied() == 0
    g.write(np.ones((1, 1, 1), np.uint16))
    g.fracture_naive(np.array([[0, 0, 0, 5]], np.uint32), MAN)
    assert g.read()[0, 0, 0] == 5
    e = vf.Grid((33, 17, 9), *BOX)
    e.voxelize(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    assert e.count_occupied() == 0
    e.fracture_flood(np.array([[1, 1, 1, 2]], np.uint32), MAN)
    out = e.read()
    assert out[1, 1, 1] == 2 and out.sum() == 2
