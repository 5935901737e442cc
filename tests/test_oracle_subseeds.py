"""Pins the oracle's sub-seed flood (FloodFracturer's prefix merge, disjoint
set and re-flood), Seeder::mergeSeeds and Seeder::nearSeeds against closed
forms, an independent pure-Python restatement of the two-phase flood on small
grids, and invariants.  CPU only."""
import heapq

import numpy as np
import pytest
from scipy import ndimage

import oracle as O

EUC, MAN, CHE = 0, 1, 2
NB6 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
NB26 = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]


def bar_seeds(*entries):
    return np.array([[x, 0, 0, lab] for x, lab in entries], np.uint32)


def test_subseed_bar_frees_the_larger_id_component():
    """Fragment 2 owns x 0..5 (id 1) and x 16..30 (id 2), split by fragment 3
    (x 6..15, ties to the lower seed index).  The id-2 part is freed and
    re-flooded from its kept neighbour, fragment 3."""
    g = np.ones((31, 1, 1), np.uint16)
    seeds = bar_seeds((0, 2 | 1 << 8), (10, 3), (20, 2 | 2 << 8))
    out = O.flood(g, seeds, 6)[:, 0, 0]
    np.testing.assert_array_equal(out[:6], 2)
    np.testing.assert_array_equal(out[6:], 3)
    # swapping the ids frees the left part instead; fragment 3 takes it from x = 6
    seeds2 = bar_seeds((0, 2 | 2 << 8), (10, 3), (20, 2 | 1 << 8))
    out2 = O.flood(g, seeds2, 6)[:, 0, 0]
    np.testing.assert_array_equal(out2[:16], 3)
    np.testing.assert_array_equal(out2[16:], 2)


def test_subseed_bar_connected_component_is_kept_whole():
    g = np.ones((31, 1, 1), np.uint16)
    seeds = bar_seeds((0, 2 | 1 << 8), (10, 2 | 3 << 8), (20, 2 | 2 << 8))
    np.testing.assert_array_equal(O.flood(g, seeds, 6), 2)


def test_subseed_freed_region_without_kept_neighbour_stays_free():
    """Two disconnected pieces; the right piece holds only fragment 2's id-2
    seed, so once freed nothing re-floods it (FREE, as unreached cells)."""
    g = np.ones((21, 1, 1), np.uint16)
    g[10] = 0
    seeds = bar_seeds((0, 2 | 1 << 8), (15, 2 | 2 << 8))
    out = O.flood(g, seeds, 6)[:, 0, 0]
    np.testing.assert_array_equal(out[:10], 2)
    assert out[10] == 0
    np.testing.assert_array_equal(out[11:], 1)


def rand_occ(shape, p, seed):
    return (np.random.default_rng(seed).random(shape) < p).astype(np.uint8)


def rand_sub_seeds(occ, n, nfrag, seed, same_id=False):
    rng = np.random.default_rng(seed)
    cells = np.argwhere(occ > 0)
    pick = cells[rng.choice(len(cells), size=n, replace=False)]
    frag = 2 + rng.integers(0, nfrag, size=n)
    sub = np.ones(n, np.int64) if same_id else rng.permutation(n) + 1
    return np.concatenate([pick, (frag | (sub << 8))[:, None]], 1).astype(np.uint32)


@pytest.mark.parametrize("nb", [6, 26])
@pytest.mark.parametrize("amp", [0, 2])
def test_subseeds_with_one_id_per_fragment_equal_plain_flood(nb, amp):
    occ = rand_occ((12, 11, 10), 0.6, 3)
    g = O.occ_to_grid(occ)
    seeds = rand_sub_seeds(occ, 9, 4, 5, same_id=True)
    plain = seeds.copy()
    plain[:, 3] &= 0xFF
    np.testing.assert_array_equal(O.flood(g, seeds, nb, 17, amp), O.flood(g, plain, nb, 17, amp))


def two_phase_python(g, seeds, nb, noise_seed, amp):
    """Independent restatement: first growth from the per-seed Dijkstra brute
    force (owners via unique labels), fragment components by scipy, then a
    heap Dijkstra from every kept cell in pure Python."""
    S = seeds.shape[0]
    uniq = seeds.copy()
    uniq[:, 3] = 2 + np.arange(S)
    first = O.flood(g, uniq, nb, noise_seed, amp, dijkstra=True).astype(np.int64)
    owner = np.where(first >= 2, first - 2, -1)
    frag = np.where(owner >= 0, seeds[np.maximum(owner, 0), 3] & 0xFF, 0)
    sub = np.where(owner >= 0, seeds[np.maximum(owner, 0), 3] >> 8, 0)
    st = ndimage.generate_binary_structure(3, 1 if nb == 6 else 3)
    keep = np.zeros(g.shape, bool)
    comps = {}
    for f in np.unique(frag[owner >= 0]):
        lab, n = ndimage.label(frag == f, structure=st)
        mins = ndimage.minimum(sub, lab, index=np.arange(1, n + 1))
        comps[f] = (lab, np.asarray(mins))
    for f, (lab, mins) in comps.items():
        m = mins.min()
        keep |= (lab > 0) & (mins[np.maximum(lab - 1, 0)] == m)
    freed = (owner >= 0) & ~keep
    if freed.any():
        occ = (g != 0) | (first >= 1)
        for s in seeds:
            occ[tuple(s[:3])] = True
        dist = {}
        heap = []
        for c in map(tuple, np.argwhere(keep)):
            dist[c] = (0, int(owner[c]))
            heap.append((0, int(owner[c]), c))
        heapq.heapify(heap)
        while heap:
            d, s, c = heapq.heappop(heap)
            if dist.get(c) != (d, s):
                continue
            for o in (NB6 if nb == 6 else NB26):
                u = (c[0] + o[0], c[1] + o[1], c[2] + o[2])
                if min(u) < 0 or u[0] >= g.shape[0] or u[1] >= g.shape[1] or u[2] >= g.shape[2] or not occ[u]:
                    continue
                lin = int(np.ravel_multi_index(u, g.shape))
                cand = (d + 1 + O.noise_delta(noise_seed, lin, amp), s)
                if u not in dist or cand < dist[u]:
                    dist[u] = cand
                    heapq.heappush(heap, (cand[0], cand[1], u))
        owner = np.full(g.shape, -1, np.int64)
        for c, (_, s) in dist.items():
            owner[c] = s
    out = np.where(owner >= 0, seeds[np.maximum(owner, 0), 3] & 0xFF, 0).astype(np.uint16)
    occ_all = (g != 0)
    for s in seeds:
        occ_all[tuple(s[:3])] = True
    out[(owner < 0) & occ_all] = 1
    return out, bool(freed.any())


@pytest.mark.parametrize("nb,amp,seed", [(6, 0, 1), (26, 0, 2), (6, 2, 3), (26, 1, 4), (6, 1, 5), (26, 2, 6)])
def test_subseed_flood_matches_python_two_phase(nb, amp, seed):
    occ = rand_occ((10, 9, 8), 0.62, seed)
    g = O.occ_to_grid(occ)
    seeds = rand_sub_seeds(occ, 12, 3, seed + 40)
    ref, freed = two_phase_python(g, seeds, nb, seed * 31, amp)
    np.testing.assert_array_equal(O.flood(g, seeds, nb, seed * 31, amp), ref)


def test_subseed_cases_exercise_the_freeing_pass():
    hits = 0
    for nb, amp, seed in [(6, 0, 1), (26, 0, 2), (6, 2, 3), (26, 1, 4), (6, 1, 5), (26, 2, 6)]:
        occ = rand_occ((10, 9, 8), 0.62, seed)
        _, freed = two_phase_python(O.occ_to_grid(occ), rand_sub_seeds(occ, 12, 3, seed + 40), nb, seed * 31, amp)
        hits += freed
    assert hits >= 3


@pytest.mark.parametrize("nb", [6, 26])
def test_subseed_fragments_connected_when_ids_distinct(nb):
    occ = rand_occ((14, 13, 12), 0.7, 8)
    g = O.occ_to_grid(occ)
    seeds = rand_sub_seeds(occ, 16, 4, 9)
    out = O.flood(g, seeds, nb, 5, 2)
    st = ndimage.generate_binary_structure(3, 1 if nb == 6 else 3)
    for f in np.unique(out[out >= 2]):
        _, n = ndimage.label(out == f, structure=st)
        assert n == 1, f"fragment {f} split into {n}"


def test_subseed_label_validation():
    g = np.ones((4, 4, 4), np.uint16)
    with pytest.raises(ValueError):
        O.flood(g, np.array([[0, 0, 0, 1 | 1 << 8]], np.uint32), 6)  # fragment id 1 is FREE
    with pytest.raises(ValueError):
        O.flood(g, np.array([[0, 0, 0, 0x10002]], np.uint32), 6)  # beyond 16 bits
    with pytest.raises(ValueError):  # the per-seed brute force takes plain labels only
        O.flood(g, np.array([[0, 0, 0, 2 | 1 << 8]], np.uint32), 6, dijkstra=True)


# ----------------------------------------------------------------- mergeSeeds
def test_merge_seeds_known_answer():
    frags = np.array([[0, 0, 0, 2], [10, 0, 0, 3]], np.uint32)
    seeds = np.array([[0, 0, 0, 2], [10, 0, 0, 3], [2, 0, 0, 4], [9, 0, 0, 5], [5, 0, 0, 6]], np.uint32)
    out = O.seed_merge(frags, seeds, EUC)
    np.testing.assert_array_equal(out[:, 3], [2 | 1 << 8, 3 | 1 << 8, 2 | 2 << 8, 3 | 2 << 8, 2 | 3 << 8])
    np.testing.assert_array_equal(out[:, :3], seeds[:, :3])


def merge_float_reference(frags, seeds, dist):
    """Seeder::mergeSeeds with the reference's float distances (glm::distance in
    float32, abs sums, max) and first-wins strict comparison."""
    f = frags[:, :3].astype(np.float32)
    out = seeds.copy()
    count = {}
    for i, s in enumerate(seeds[:, :3].astype(np.float32)):
        d = f - s
        if dist == EUC:
            dd = np.sqrt((d * d).sum(axis=1, dtype=np.float32)).astype(np.float32)
        elif dist == MAN:
            dd = np.abs(d).sum(axis=1, dtype=np.float32)
        else:
            dd = np.abs(d).max(axis=1)
        best, near = np.float32(np.finfo(np.float32).max), -1
        for k, v in enumerate(dd):
            if v < best:
                best, near = v, k
        w = int(frags[near, 3])
        count[w] = count.get(w, 0) + 1
        out[i, 3] = w | (count[w] << 8)
    return out


@pytest.mark.parametrize("dist", [EUC, MAN, CHE])
def test_merge_seeds_matches_float_reference(dist):
    rng = np.random.default_rng(dist + 3)
    frags = np.concatenate([rng.integers(0, 300, (12, 3)), (2 + np.arange(12))[:, None]], 1).astype(np.uint32)
    seeds = np.concatenate([frags, np.concatenate([rng.integers(0, 300, (60, 3)), np.zeros((60, 1), int)], 1)])
    seeds = seeds.astype(np.uint32)
    np.testing.assert_array_equal(O.seed_merge(frags, seeds, dist), merge_float_reference(frags, seeds, dist))


def test_merge_seeds_sub_id_limit():
    frags = np.array([[0, 0, 0, 2]], np.uint32)
    seeds = np.zeros((256, 4), np.uint32)
    with pytest.raises(ValueError):
        O.seed_merge(frags, seeds, EUC)
    O.seed_merge(frags, seeds[:255], EUC)


# ------------------------------------------------------------------ nearSeeds
def shell_grid(n, wall):
    """A hollow box: occupied shell of the given thickness, empty inside."""
    g = np.zeros((n, n, n), np.uint16)
    g[:] = 1
    g[wall:n - wall, wall:n - wall, wall:n - wall] = 0
    return g


@pytest.mark.parametrize("impacts,num_seeds,spreading,key", [(1, 32, 5, 3), (3, 32, 5, 4), (4, 12, 2, 5)])
def test_near_seeds_properties(impacts, num_seeds, spreading, key):
    g = shell_grid(40, 3)
    frags = O.seed_uniform(g, 6, 11, 1)
    out = O.seed_near(g, frags, impacts, num_seeds, spreading, key)
    np.testing.assert_array_equal(out[:6], frags)
    new = out[6:]
    assert 1 <= len(new) <= num_seeds
    cells = [tuple(c) for c in new[:, :3]]
    assert len(set(cells)) == len(cells)
    lin = [np.ravel_multi_index(c, g.shape) for c in cells]
    assert lin == sorted(lin)
    np.testing.assert_array_equal(new[:, 3], frags[-1, 3] + 1 + np.arange(len(new)))
    empty_near = ndimage.maximum_filter((g == 0).astype(np.uint8), size=3, mode="nearest") > 0
    mind = 40 // 2
    for c in cells:
        assert g[c] != 0 and empty_near[c]
        d2 = ((frags[:, :3].astype(np.int64) - np.array(c)) ** 2).sum(axis=1)
        assert d2.min() <= mind * mind
    np.testing.assert_array_equal(out, O.seed_near(g, frags, impacts, num_seeds, spreading, key))
    if impacts == 1:  # one impact draws 1 .. num_seeds - 1 seeds, all around one fragment seed
        assert len(new) < num_seeds


def test_near_seeds_rejects_bad_spreading():
    g = shell_grid(8, 2)
    frags = np.array([[0, 0, 0, 2]], np.uint32)
    with pytest.raises(ValueError):
        O.seed_near(g, frags, 1, 4, 9, 1)  # spreading above an edge: the biased draw divides by zero
    with pytest.raises(ValueError):
        O.seed_near(g, frags[:0], 1, 4, 2, 1)
