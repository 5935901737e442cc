"""Pins the oracle's erosion (RegularGrid::erode, erodeGrid-comp.glsl,
removeIsolatedRegionsGrid-comp.glsl) against independent scipy/numpy
computations: window counts by correlation, FP32 activation tests, closed
forms on a solid box, mask shapes.  CPU only."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O

SQUARE, ELLIPSE, CROSS = 0, 1, 2


def test_masks():
    m, a = O.erosion_mask(SQUARE, 3)
    assert m.shape == (3, 3, 3) and m.all() and a == np.float32(1.0)
    m, a = O.erosion_mask(SQUARE, 4)  # even sizes grow by one
    assert m.shape == (5, 5, 5) and m.all()
    m, a = O.erosion_mask(CROSS, 5)
    assert m.sum() == 3 * 5 - 2
    assert a == np.float32(np.float32(1.0 / 3.0) * np.float32(125)) / np.float32(125)
    # ELLIPSE: float distance below centre + FLT_EPSILON; for side 3 the epsilon admits d = 1 (7 cells),
    # for side 5 it rounds away (2 + eps == 2 in FP32) so d = 2 is excluded (27 cells)
    for side, cells in [(3, 7), (5, 27), (7, 123)]:
        m, a = O.erosion_mask(ELLIPSE, side)
        c = side // 2
        idx = np.indices((side,) * 3).reshape(3, -1).T.astype(np.float32) - np.float32(c)
        d = np.sqrt((idx * idx).sum(axis=1, dtype=np.float32)).astype(np.float32)
        ref = d < np.float32(c) + np.finfo(np.float32).eps
        np.testing.assert_array_equal(m.reshape(-1).astype(bool), ref)
        assert m.sum() == ref.sum()
        assert a == np.float32(ref.sum()) / np.float32(side ** 3)
    assert O.erosion_mask(ELLIPSE, 3)[0].sum() == 7 and O.erosion_mask(ELLIPSE, 5)[0].sum() == 27


def test_noise_uniform():
    u = np.array([O.erosion_noise(9, c) for c in range(20000)], np.float32)
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert abs((u < 0.3).mean() - 0.3) < 0.015


def erode_step_scipy(grid, mask, limit, probability, noise_seed):
    """erodeGrid-comp.glsl restated with correlations: per distinct value v,
    count = correlate(grid == v, mask) (out-of-grid cells count 0), global =
    in-bounds window cells; FP32 division and comparison."""
    side = mask.shape[0]
    noise = np.array([O.erosion_noise(noise_seed, c) for c in range(grid.size)], np.float32).reshape(grid.shape)
    glob = ndimage.correlate(np.ones(grid.shape, np.int64), np.ones((side,) * 3, np.int64), mode="constant", cval=0)
    out = grid.copy()
    eligible = (grid > 1) & ((grid & 0x7FFF) != 0) & (noise < np.float32(probability))
    for v in np.unique(grid[eligible]):
        cnt = ndimage.correlate((grid == v).astype(np.int64), mask.astype(np.int64), mode="constant", cval=0)
        act = cnt.astype(np.float32) / glob.astype(np.float32)
        out[eligible & (grid == v) & (act < np.float32(limit))] = 0
    return out


def labelled_grid(shape, seed, p=0.8, labels=4):
    rng = np.random.default_rng(seed)
    g = np.where(rng.random(shape) < p, rng.integers(2, 2 + labels, size=shape), 0).astype(np.uint16)
    g[rng.random(shape) < 0.3] |= 0x8000
    g[rng.random(shape) < 0.05] = 1
    return g


@pytest.mark.parametrize("kind,size", [(SQUARE, 3), (ELLIPSE, 3), (CROSS, 3), (ELLIPSE, 5), (SQUARE, 4)])
@pytest.mark.parametrize("prob,thr", [(1.0, 0.5), (0.5, 0.8)])
def test_erode_step_matches_scipy(kind, size, prob, thr):
    g = labelled_grid((11, 9, 10), size * 7 + kind)
    mask, act = O.erosion_mask(kind, size)
    limit = np.float32(act * np.float32(thr))
    np.testing.assert_array_equal(O.erode_step(g, mask, limit, prob, 5), erode_step_scipy(g, mask, limit, prob, 5))


def test_erode_step_box_closed_form():
    """One label filling a box inside EMPTY space, SQUARE 3, p = 1: a face cell
    sees 18/27 of its window, an edge cell 12/27, a corner 8/27, the interior
    27/27.  Threshold 0.5 removes edges and corners, 0.7 the faces too."""
    g = np.zeros((9, 9, 9), np.uint16)
    g[2:7, 2:7, 2:7] = 5
    mask, act = O.erosion_mask(SQUARE, 3)
    out = O.erode_step(g, mask, act * np.float32(0.5), 1.0, 1)
    box = np.zeros_like(g, bool)
    box[2:7, 2:7, 2:7] = True
    coords = np.indices(g.shape)
    extreme = sum(((coords[a] == 2) | (coords[a] == 6)).astype(int) for a in range(3)) * box
    np.testing.assert_array_equal(out == 5, box & (extreme <= 1))
    out7 = O.erode_step(g, mask, act * np.float32(0.7), 1.0, 1)
    np.testing.assert_array_equal(out7 == 5, box & (extreme == 0))
    # threshold 0 or probability 0 erode nothing
    np.testing.assert_array_equal(O.erode_step(g, mask, np.float32(0), 1.0, 1), g)
    np.testing.assert_array_equal(O.erode_step(g, mask, act, 0.0, 1), g)


def test_remove_isolated_cells_matches_scipy():
    g = labelled_grid((12, 10, 11), 3, p=0.9, labels=2)
    ones = np.ones((3, 3, 3), np.int64)
    ref = g.copy()
    for v in np.unique(g):
        same = (g == v).astype(np.int64)
        cnt = ndimage.correlate(same, ones, mode="constant", cval=0) - 1
        ref[(g == v) & (cnt < 6)] = 0
    np.testing.assert_array_equal(O.remove_isolated_cells(g), ref)


def test_erode_composes_boundaries_steps_and_cleanup():
    g = labelled_grid((10, 12, 9), 8, p=0.95, labels=3) & 0x7FFF
    g[g == 1] = 2
    mask, act = O.erosion_mask(ELLIPSE, 3)
    limit = act * np.float32(0.6)
    ref = g.copy()
    for _ in range(3):
        ref = erode_step_scipy(O.detect_boundaries(ref, 1), mask, limit, 0.5, 77)
    ref = O.remove_isolated_cells(ref)
    np.testing.assert_array_equal(O.erode(g, ELLIPSE, 3, 3, 0.5, 0.6, 77), ref)
