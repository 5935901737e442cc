"""GPU parity of RegularGrid::erode and the grid's isolated-cell pass through
the C ABI against the CPU oracle.  Integer outputs must be bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EUC, MAN, CHE = 0, 1, 2
SQUARE, ELLIPSE, CROSS = 0, 1, 2
BOX = ([-1.03, -1.03, -1.03], [1.03, 1.03, 1.03])


def fractured(vf, dims, S, seed):
    xyz, tri = vf.vessel_mesh(seed, *vf.vessel_segments(20000))
    g = vf.Grid(dims, *BOX)
    g.voxelize(xyz, tri)
    base = g.read()
    seeds = O.seed_uniform(base, S, seed)
    g.fracture_naive(seeds, EUC)
    return g, O.naive(base, seeds, EUC)


@pytest.mark.parametrize("kind,size,iters,prob,thr", [
    (ELLIPSE, 3, 3, 0.5, 0.5),   # FractureParameters defaults
    (SQUARE, 3, 2, 1.0, 0.6),
    (CROSS, 5, 1, 0.7, 0.9),
    (ELLIPSE, 5, 2, 0.9, 0.7),
    (SQUARE, 4, 1, 0.5, 0.8),    # even size grows to 5
    (ELLIPSE, 7, 1, 1.0, 0.5),
])
def test_erode_matches_oracle(vf, kind, size, iters, prob, thr):
    print("erode", kind, size, iters, prob, thr, flush=True)
    g, lab = fractured(vf, (64, 60, 56), 8, 3)
    np.testing.assert_array_equal(g.read(), lab)
    g.erode(kind, size, iters, prob, thr, noise_seed=41)
    ref = O.erode(lab, kind, size, iters, prob, thr, 41)
    out = g.read()
    np.testing.assert_array_equal(out, ref)
    assert (out == 0).sum() > (lab == 0).sum()  # something was eroded
    np.testing.assert_array_equal(g.occupancy(), (ref != 0).astype(np.uint8))
    g.close()


def test_erode_ragged_and_masked_input(vf):
    rng = np.random.default_rng(5)
    dims = (33, 47, 29)
    grid = np.where(rng.random(dims) < 0.85, rng.integers(2, 6, size=dims), 0).astype(np.uint16)
    grid[rng.random(dims) < 0.2] |= 0x8000
    g = vf.Grid(dims, *BOX)
    g.write(grid)
    g.erode(ELLIPSE, 3, 4, 0.6, 0.7, noise_seed=3)
    np.testing.assert_array_equal(g.read(), O.erode(grid, ELLIPSE, 3, 4, 0.6, 0.7, 3))
    with pytest.raises(vf.VfError):
        g.erode(ELLIPSE, 17, 1, 0.5, 0.5)
    with pytest.raises(vf.VfError):
        g.erode(7, 3, 1, 0.5, 0.5)
    g.close()


def test_remove_isolated_cells_matches_oracle(vf):
    rng = np.random.default_rng(8)
    dims = (40, 38, 36)
    grid = np.where(rng.random(dims) < 0.9, rng.integers(2, 4, size=dims), 0).astype(np.uint16)
    g = vf.Grid(dims, *BOX)
    g.write(grid)
    g.remove_isolated_cells()
    np.testing.assert_array_equal(g.read(), O.remove_isolated_cells(grid))
    g.close()
