"""vf_fracture_naive_noalloc, the entry point that allocates nothing, against
vf_fracture_naive on the same cells and seeds, through the C ABI: two calls
in a row on one GPU, the second on a smaller grid (the scratch array keeps
the first call's occupancy bytes, which must not leak into the second), for
every distance; the size limit and bad arguments are rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

BOX = ([-1.03] * 3, [1.03] * 3)


def _device_ptr(vf, g):
    p = C.c_void_p()
    assert vf.lib().vf_grid_device_ptr(g._h, C.byref(p)) == 0
    return p.value


def _noalloc(vf, dev_ptr, dims, seeds, dist):
    lib = vf.lib()
    fn = lib.vf_fracture_naive_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]
    d = np.asarray(dims, np.uint32)
    s = np.ascontiguousarray(seeds, np.uint32)
    return fn(dev_ptr, d.ctypes.data, s.ctypes.data if len(s) else None, len(s), dist, None)


def _case(vf, dims, S, dist, seed):
    xyz, tri = vf.vessel_mesh(seed, *vf.vessel_segments(9000))
    with vf.Grid(dims, *BOX) as g:
        g.voxelize(xyz, tri)
        base = g.read()
        seeds = g.seed_uniform(S, seed)
        g.fracture_naive(seeds, dist)
        want = g.read()
    with vf.Grid(dims, *BOX) as h:
        h.write(base)
        ptr = _device_ptr(vf, h)
        assert _noalloc(vf, ptr, dims, seeds, dist) == 0, vf.lib().vf_last_error()
        got = h.read()
    return got, want, base, seeds


@pytest.mark.parametrize("dist", [0, 1, 2])
def test_noalloc_naive_equals_naive_twice_smaller_second(vf, dist):
    got1, want1, base1, seeds1 = _case(vf, (96, 88, 104), 23, dist, 5 + dist)
    np.testing.assert_array_equal(got1, want1)
    got2, want2, base2, seeds2 = _case(vf, (40, 36, 44), 7, dist, 9 + dist)
    np.testing.assert_array_equal(got2, want2)
    np.testing.assert_array_equal(got2, O.naive(base2, seeds2, dist))


def test_noalloc_rejects_beyond_limit_and_bad_arguments(vf):
    lib = vf.lib()
    limit = 1 << 24
    with vf.Grid((8, 8, 8), *BOX) as g:
        ptr = _device_ptr(vf, g)
        seeds = np.array([[1, 1, 1, 2]], np.uint32)
        assert _noalloc(vf, ptr, (257, 256, 256), seeds, 0) == -1  # more cells than the scratch holds
        assert 257 * 256 * 256 > limit
        assert _noalloc(vf, ptr, (8, 8, 8), np.array([[8, 0, 0, 2]], np.uint32), 0) == -1  # seed outside
        assert _noalloc(vf, ptr, (8, 8, 8), seeds, 7) == -1  # unknown distance
        assert _noalloc(vf, ptr + 2, (8, 8, 8), seeds, 0) == -1  # misaligned cells
        assert _noalloc(vf, None, (8, 8, 8), seeds, 0) == -1
        g.write(np.ones((8, 8, 8), np.uint16))
        assert _noalloc(vf, ptr, (8, 8, 8), seeds, 1) == 0
        assert np.all(g.read() == 2)
    assert lib.vf_fracture_naive_noalloc is not None
