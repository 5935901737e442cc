"""GPU parity of the Seeder extras and the flood's sub-seed pass, through the
C ABI, against the CPU oracle: merged seeds (Seeder::mergeSeeds labels) in the
flood, impact seeds (Seeder::nearSeeds), and the CADScene::fractureModel
sequence with extra seeds.  Integer outputs must be bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EUC, MAN, CHE = 0, 1, 2
BOX = ([-1.03, -1.03, -1.03], [1.03, 1.03, 1.03])


def voxelized(vf, dims, seed, tris=20000):
    xyz, tri = vf.vessel_mesh(seed, *vf.vessel_segments(tris))
    g = vf.Grid(dims, *BOX)
    g.voxelize(xyz, tri)
    return g


def test_subseed_bar_closed_form(vf):
    g = vf.Grid((31, 1, 1), *BOX)
    g.write(np.ones((31, 1, 1), np.uint16))
    seeds = np.array([[0, 0, 0, 2 | 1 << 8], [10, 0, 0, 3], [20, 0, 0, 2 | 2 << 8]], np.uint32)
    g.fracture_flood(seeds, MAN)
    out = g.read()[:, 0, 0]
    np.testing.assert_array_equal(out[:6], 2)
    np.testing.assert_array_equal(out[6:], 3)
    g.close()


@pytest.mark.parametrize("nb,amp,dims", [(6, 0, (40, 36, 44)), (26, 0, (48, 48, 48)), (6, 2, (37, 53, 29)),
                                         (26, 2, (64, 64, 64))])
def test_subseed_flood_random_grid(vf, nb, amp, dims):
    print("subseed random", nb, amp, dims, flush=True)
    rng = np.random.default_rng(sum(dims) + nb)
    base = (rng.random(dims) < 0.7).astype(np.uint16)
    cells = np.argwhere(base > 0)
    pick = cells[rng.choice(len(cells), size=40, replace=False)]
    frag = 2 + rng.integers(0, 6, size=40)
    sub = rng.permutation(40) + 1
    seeds = np.concatenate([pick, (frag | (sub << 8))[:, None]], 1).astype(np.uint32)
    g = vf.Grid(dims, *BOX)
    g.write(base)
    g.fracture_flood(seeds, CHE, nb, noise_seed=77, noise_amp=amp)
    out = g.read()
    np.testing.assert_array_equal(out, O.flood(base, seeds, nb, 77, amp))
    plain = seeds.copy()
    plain[:, 3] &= 0xFF
    assert (out != O.flood(base, plain, nb, 77, amp)).sum() > 1000  # the freeing pass changed the result
    g.close()


@pytest.mark.parametrize("algo,dist,dims,S", [("flood", EUC, (72, 72, 72), 6), ("flood", MAN, (64, 56, 60), 9),
                                              ("naive", CHE, (64, 64, 64), 5), ("flood", CHE, (96, 96, 96), 12)])
def test_fracture_model_with_extra_seeds(vf, algo, dist, dims, S):
    """CADScene::fractureModel (CADScene.cpp:624-691) with numExtraSeeds = 2 S:
    OUTER fragment seeds, BOTH extra seeds (another key) prefixed by the
    fragment seeds, mergeSeeds, then the fragment seeds followed by the merged
    list."""
    print("fractureModel extra seeds", algo, dist, dims, S, flush=True)
    g = voxelized(vf, dims, seed=S + 2)
    base = g.read()
    frags = g.seed_uniform(S, 31, vf.SEED_OUTER)
    extra = g.seed_uniform(2 * S, 32, vf.SEED_BOTH)
    np.testing.assert_array_equal(frags, O.seed_uniform(base, S, 31, 1))
    extra = np.concatenate([frags, extra])
    merged = vf.seed_merge(frags, extra, EUC)
    np.testing.assert_array_equal(merged, O.seed_merge(frags, extra, EUC))
    seeds = np.concatenate([frags, merged])
    if algo == "flood":
        g.fracture_flood(seeds, dist, noise_seed=5, noise_amp=1)
        nb = 6 if dist == MAN else 26
        ref = O.flood(base, seeds, nb, 5, 1)
    else:
        g.fracture_naive(seeds, dist)
        ref = O.naive(base, seeds, dist)
    np.testing.assert_array_equal(g.read(), ref)
    g.close()


@pytest.mark.parametrize("impacts,num_seeds,spreading,key", [(1, 32, 5, 1), (3, 32, 5, 2), (5, 40, 3, 3)])
def test_seed_near_matches_oracle(vf, impacts, num_seeds, spreading, key):
    print("nearSeeds", impacts, num_seeds, spreading, flush=True)
    g = voxelized(vf, (80, 80, 80), seed=4)
    base = g.read()
    frags = g.seed_uniform(8, 9, vf.SEED_OUTER)
    out = g.seed_near(frags, impacts, num_seeds, spreading, key)
    np.testing.assert_array_equal(out, O.seed_near(base, frags, impacts, num_seeds, spreading, key))
    assert out.shape[0] > frags.shape[0]
    with pytest.raises(vf.VfError):
        g.seed_near(frags, 1, 8, 0, key)
    g.close()
