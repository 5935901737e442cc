"""Pins the oracle's fracture functions (naive, flood, isolated regions, seeding)
against independent computations: numpy brute force with the reference's own
float distances, closed forms on solid boxes, per-seed Dijkstra, invariants.
CPU only."""
import numpy as np
import pytest

import oracle as O

EUC, MAN, CHE = 0, 1, 2


def random_occupancy(shape, p, seed):
    rng = np.random.default_rng(seed)
    return (rng.random(shape) < p).astype(np.uint8)


def random_seeds(occ, n, seed):
    rng = np.random.default_rng(seed)
    cells = np.argwhere(occ > 0)
    pick = cells[rng.choice(len(cells), size=n, replace=False)]
    pick = pick[np.lexsort((pick[:, 2], pick[:, 1], pick[:, 0]))]
    return np.concatenate([pick, (2 + np.arange(n))[:, None]], axis=1).astype(np.uint32)


def naive_numpy_float(grid, seeds, dist):
    """NaiveFracturer::buildCPU with the reference's float distances (sqrt for
    Euclidean) and first-index argmin, vectorised with numpy."""
    idx = np.indices(grid.shape).reshape(3, -1).T.astype(np.float64)
    s = seeds[:, :3].astype(np.float64)
    d = idx[:, None, :] - s[None, :, :]
    if dist == EUC:
        D = np.sqrt((d * d).sum(-1)).astype(np.float32)
    elif dist == MAN:
        D = np.abs(d).sum(-1).astype(np.float32)
    else:
        D = np.abs(d).max(-1).astype(np.float32)
    best = np.argmin(D, axis=1)  # first minimum = strict '<' in list order
    out = grid.ravel().copy()
    occ = out != 0
    out[occ] = seeds[best[occ], 3]
    return out.reshape(grid.shape)


@pytest.mark.parametrize("dist", [EUC, MAN, CHE])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_naive_matches_float_reference_semantics(dist, seed):
    occ = random_occupancy((9, 11, 13), 0.5, seed)
    g = O.occ_to_grid(occ)
    seeds = random_seeds(occ, 7, seed + 10)
    got = O.naive(g, seeds, dist)
    np.testing.assert_array_equal(got, naive_numpy_float(g, seeds, dist))


def naive_float32_cells(cells, seeds, dist):
    """Labels of the given cells by the reference's float arithmetic carried out in float32
    (coordinate differences, squares summed in order, sqrt), first minimum wins."""
    c = cells.astype(np.float32)
    s = seeds[:, :3].astype(np.float32)
    d = c[:, None, :] - s[None, :, :]
    if dist == EUC:
        D = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
    elif dist == MAN:
        D = np.abs(d[..., 0]) + np.abs(d[..., 1]) + np.abs(d[..., 2])
    else:
        D = np.maximum(np.abs(d[..., 0]), np.maximum(np.abs(d[..., 1]), np.abs(d[..., 2])))
    assert D.dtype == np.float32
    return seeds[np.argmin(D, axis=1), 3]


def test_naive_euclid_float_roots_tie_as_in_reference():
    # two squared distances one apart whose float32 roots are equal: the reference's strict '<'
    # keeps the first seed although the second is nearer in exact arithmetic
    grid = np.zeros((2048, 2048, 1), np.uint16)
    grid[0, 0, 0] = 1
    seeds = np.array([[659, 1951, 0, 2], [616, 1965, 0, 3]], np.uint32)
    d2 = [int(s[0]) ** 2 + int(s[1]) ** 2 for s in seeds]
    assert d2 == [4240682, 4240681]
    assert np.sqrt(np.float32(d2[0])) == np.sqrt(np.float32(d2[1]))
    got = O.naive(grid, seeds, EUC)
    assert got[0, 0, 0] == 2
    assert naive_float32_cells(np.array([[0, 0, 0]]), seeds, EUC)[0] == 2


@pytest.mark.parametrize("dist", [EUC, MAN, CHE])
def test_naive_large_grid_matches_float32_reference(dist):
    # a 2048 x 2048 x 2 grid: squared distances up to 8.4M, where float roots of distinct integers collide
    rng = np.random.default_rng(11)
    shape = (2048, 2048, 2)
    grid = np.zeros(shape, np.uint16)
    cells = rng.integers(0, [2048, 2048, 2], size=(4000, 3))
    grid[cells[:, 0], cells[:, 1], cells[:, 2]] = 1
    seeds = random_seeds(grid, 96, 12)
    got = O.naive(grid, seeds, dist)
    occ = np.argwhere(grid != 0)
    np.testing.assert_array_equal(got[occ[:, 0], occ[:, 1], occ[:, 2]], naive_float32_cells(occ, seeds, dist))
    assert np.count_nonzero(got) == len(occ)


@pytest.mark.parametrize("dist", [EUC, MAN, CHE])
def test_naive_invariants(dist):
    occ = random_occupancy((16, 16, 16), 0.3, 7)
    g = O.occ_to_grid(occ)
    seeds = random_seeds(occ, 12, 8)
    got = O.naive(g, seeds, dist)
    assert np.all(got[occ == 0] == 0)
    assert set(np.unique(got[occ == 1])) <= set(seeds[:, 3].tolist())
    for s in seeds:  # every seed cell carries its own label
        assert got[s[0], s[1], s[2]] == s[3]


def test_naive_tie_goes_to_first_seed():
    g = np.ones((5, 1, 1), np.uint16)
    seeds = np.array([[4, 0, 0, 7], [0, 0, 0, 9]], np.uint32)  # cell 2 is equidistant
    got = O.naive(g, seeds, MAN)
    assert got[2, 0, 0] == 7  # first in list order wins, whatever its label
    got = O.naive(g, seeds[::-1].copy(), MAN)
    assert got[2, 0, 0] == 9


def test_naive_box_halfspaces_closed_form():
    """Two seeds on the x axis of a solid box: Chebyshev/Manhattan partition is
    the half-space x <= mid (ties at the mid plane go to the first seed)."""
    g = np.ones((10, 7, 7), np.uint16)
    seeds = np.array([[2, 3, 3, 2], [7, 3, 3, 3]], np.uint32)
    for dist in (MAN, EUC):
        got = O.naive(g, seeds, dist)
        x = np.arange(10)[:, None, None]
        expect = np.where(x <= 4, 2, 3) * np.ones_like(g)
        np.testing.assert_array_equal(got, expect)


def test_naive_without_seeds_leaves_grid():
    g = O.occ_to_grid(random_occupancy((6, 6, 6), 0.5, 1))
    g[g == 1] = 5
    np.testing.assert_array_equal(O.naive(g, np.zeros((0, 4), np.uint32), MAN), g)


# ------------------------------------------------------------------ flood
@pytest.mark.parametrize("nb", [6, 26])
@pytest.mark.parametrize("amp", [0, 1, 2])
@pytest.mark.parametrize("seed", [11, 12])
def test_flood_matches_dijkstra(nb, amp, seed):
    occ = random_occupancy((12, 10, 14), 0.65, seed)
    g = O.occ_to_grid(occ)
    seeds = random_seeds(occ, 6, seed + 100)
    a = O.flood(g, seeds, nb, noise_seed=seed * 7919, amp=amp)
    b = O.flood(g, seeds, nb, noise_seed=seed * 7919, amp=amp, dijkstra=True)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("nb,dist", [(6, MAN), (26, CHE)])
def test_flood_equals_naive_on_solid_box(nb, dist):
    """Inside an axis-aligned box the 6-/26-connected graph distance equals
    L1/L-inf, so zero-noise flood must equal naive cell for cell, ties included."""
    g = np.ones((11, 9, 13), np.uint16)
    rng = np.random.default_rng(5)
    cells = rng.choice(g.size, size=9, replace=False)
    cells.sort()
    xyz = np.stack(np.unravel_index(cells, g.shape), 1)
    seeds = np.concatenate([xyz, (2 + np.arange(9))[:, None]], 1).astype(np.uint32)
    np.testing.assert_array_equal(O.flood(g, seeds, nb), O.naive(g, seeds, dist))


def _components(mask, nb):
    from scipy import ndimage
    st = ndimage.generate_binary_structure(3, 1 if nb == 6 else 3)
    return ndimage.label(mask, structure=st)


@pytest.mark.parametrize("nb", [6, 26])
def test_flood_fragments_connected_and_unreached_free(nb):
    occ = random_occupancy((14, 14, 14), 0.55, 3)
    occ[7, :, :] = 0  # a wall splits the grid; seeds only on one side
    g = O.occ_to_grid(occ)
    left = occ.copy()
    left[7:] = 0
    seeds = random_seeds(left, 5, 4)
    out = O.flood(g, seeds, nb, noise_seed=99, amp=2)
    for s in seeds:
        lab = s[3]
        comp, n = _components(out == lab, nb)
        assert n == 1, f"fragment {lab} split into {n} components"
        assert out[s[0], s[1], s[2]] == lab
    reach, _ = _components(occ > 0, nb)
    seeded = set(reach[s[0], s[1], s[2]] for s in seeds)
    unreached = (occ > 0) & ~np.isin(reach, list(seeded))
    assert np.all(out[unreached] == 1)
    assert np.all(out[occ == 0] == 0)


def test_flood_duplicate_seed_last_wins_and_label_range():
    g = np.ones((6, 6, 6), np.uint16)
    seeds = np.array([[1, 1, 1, 2], [1, 1, 1, 3], [4, 4, 4, 4]], np.uint32)
    out = O.flood(g, seeds, 6)
    assert out[1, 1, 1] == 3
    assert 2 not in np.unique(out)
    with pytest.raises(ValueError):
        O.flood(g, np.array([[0, 0, 0, 256]], np.uint32), 6)
    with pytest.raises(ValueError):
        O.flood(g, np.array([[0, 0, 9, 2]], np.uint32), 6)


def test_flood_seed_on_empty_cell_is_a_source():
    g = np.ones((5, 5, 5), np.uint16)
    g[2, 2, 2] = 0
    out = O.flood(g, np.array([[2, 2, 2, 5]], np.uint32), 6)
    assert out[2, 2, 2] == 5 and np.all(out == 5)


def test_noise_delta_range_and_uniformity():
    v = np.array([O.noise_delta(1234, i, 2) for i in range(30000)])
    assert v.min() == 0 and v.max() == 2
    frac = np.bincount(v) / v.size
    assert np.all(np.abs(frac - 1 / 3) < 0.02)
    assert all(O.noise_delta(1234, i, 0) == 0 for i in range(100))


# ------------------------------------------------------------ isolated regions
def test_remove_isolated_regions_keeps_seeded_components():
    g = np.zeros((8, 8, 8), np.uint16)
    g[0:3, 0:3, 0:3] = 2
    g[5:8, 5:8, 5:8] = 2   # same label, not connected to its seed
    g[0:3, 5:8, 0:3] = 3
    seeds = np.array([[1, 1, 1, 2], [1, 6, 1, 3]], np.uint32)
    out = O.remove_isolated(g, seeds)
    expect = g.copy()
    expect[5:8, 5:8, 5:8] = 0
    np.testing.assert_array_equal(out, expect)


def test_remove_isolated_regions_matches_label_components():
    occ = random_occupancy((12, 12, 12), 0.6, 21)
    g = O.occ_to_grid(occ)
    seeds = random_seeds(occ, 6, 22)
    lab = O.naive(g, seeds, EUC)
    out = O.remove_isolated(lab, seeds)
    expect = np.zeros_like(lab)
    for s in seeds:
        comp, _ = _components(lab == s[3], 6)
        expect[comp == comp[s[0], s[1], s[2]]] = s[3]
    np.testing.assert_array_equal(out, expect)


# -------------------------------------------------------------------- seeding
@pytest.mark.parametrize("loc", [0, 1, 2])
def test_seed_uniform_properties(loc):
    occ = random_occupancy((20, 18, 16), 0.98 if loc == 0 else 0.4, 31)  # 0 = INNER needs interior cells
    g = O.occ_to_grid(occ)
    seeds = O.seed_uniform(g, 25, 77, loc)
    cells = [tuple(s[:3]) for s in seeds]
    assert len(set(cells)) == 25
    assert all(occ[c] for c in cells)
    lin = [np.ravel_multi_index(c, occ.shape) for c in cells]
    assert lin == sorted(lin)
    assert list(seeds[:, 3]) == list(range(2, 27))
    from scipy import ndimage
    empty_near = ndimage.maximum_filter((occ == 0).astype(np.uint8), size=3, mode="nearest") > 0
    if loc == 1:  # OUTER
        assert all(empty_near[c] for c in cells)
    if loc == 0:  # INNER
        assert not any(empty_near[c] for c in cells)
    # deterministic, and the draw sequence is the documented hash -> rank map
    np.testing.assert_array_equal(seeds, O.seed_uniform(g, 25, 77, loc))


def test_seed_uniform_first_draw_rank():
    occ = np.ones((4, 4, 4), np.uint8)
    g = O.occ_to_grid(occ)
    s = O.seed_uniform(g, 1, 5, 2)
    r = (O.hash32(5, 0) * 64) >> 32
    assert np.ravel_multi_index(tuple(s[0, :3]), occ.shape) == r
    with pytest.raises(ValueError):
        O.seed_uniform(g, 65, 5, 2)


def test_hash32_numpy_twin_matches_c():
    keys = np.array([0, 1, 2**40 + 7, 123456789], np.uint64)
    for k in keys:
        cs = np.arange(50, dtype=np.uint32)
        np.testing.assert_array_equal(O._hash32_np(np.full(50, k, np.uint64), cs),
                                      np.array([O.hash32(int(k), int(c)) for c in cs], np.uint32))
