"""ctypes binding of the CPU oracle (oracle/libvf_oracle.so) for the tests.

The oracle is the checker only: nothing in the product imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"
# VFO_ORACLE_LIB selects another build of the same sources (the sanitizer build)
LIB_PATH = Path(os.environ.get("VFO_ORACLE_LIB", ORACLE_DIR / "libvf_oracle.so"))

_lib = None


class Mesh(C.Structure):
    _fields_ = [
        ("nv", C.c_uint32), ("nt", C.c_uint32), ("nl", C.c_uint32),
        ("xyz", C.POINTER(C.c_float)), ("xyz64", C.POINTER(C.c_double)), ("edge", C.POINTER(C.c_uint64)),
        ("tri", C.POINTER(C.c_uint32)), ("tri_flag", C.POINTER(C.c_uint8)), ("tri_cube", C.POINTER(C.c_uint64)),
        ("labels", C.POINTER(C.c_uint16)), ("voff", C.POINTER(C.c_uint32)), ("toff", C.POINTER(C.c_uint32)),
    ]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
    L = C.CDLL(str(LIB_PATH))
    P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    sig = {
        "vfo_sat_triangle_box": (I, [P, P, P, P, P]),
        "vfo_voxelize": (C.c_int64, [P, U32, P, U32, P, P, P, P, P, P]),
        "vfo_occ_to_grid": (None, [P, P, C.c_size_t]),
        "vfo_homogenize": (None, [P, C.c_size_t]),
        "vfo_undo_mask": (None, [P, C.c_size_t]),
        "vfo_unmask_rightmost": (None, [P, C.c_size_t, U32]),
        "vfo_detect_boundaries": (None, [P, P, P, I]),
        "vfo_histogram": (None, [P, C.c_size_t, P]),
        "vfo_naive": (None, [P, P, P, P, U32, I]),
        "vfo_noise_delta": (U32, [U64, U32, U32]),
        "vfo_hash32": (U32, [U64, U32]),
        "vfo_flood": (I, [P, P, P, P, U32, I, U64, U32]),
        "vfo_flood_dijkstra": (I, [P, P, P, P, U32, I, U64, U32]),
        "vfo_remove_isolated": (None, [P, P, P, P, U32]),
        "vfo_seed_uniform": (I, [P, P, U32, U64, I, P]),
        "vfo_seed_merge": (I, [P, U32, P, U32, I]),
        "vfo_erosion_mask": (I, [I, U32, P, C.POINTER(C.c_float)]),
        "vfo_erosion_noise": (C.c_float, [U64, U32]),
        "vfo_erode_step": (None, [P, P, P, P, U32, C.c_float, C.c_float, U64]),
        "vfo_remove_isolated_cells": (None, [P, P, P]),
        "vfo_erode": (I, [P, P, I, U32, U32, C.c_float, C.c_float, U64]),
        "vfo_near_key": (U64, [U64, U32]),
        "vfo_seed_near": (I, [P, P, P, U32, U32, U32, U32, U64, P, P]),
        "vfo_marching_cubes": (I, [P, P, P, P, U32, C.POINTER(Mesh)]),
        "vfo_mesh_free": (None, [C.POINTER(Mesh)]),
        "vfo_mesh_smooth": (None, [C.POINTER(Mesh), U32, C.c_float, U32, C.c_float]),
        "vfo_mc_packed": (U64, [U32]),
        "vfo_rle_encode": (C.c_size_t, [P, P, P, C.c_size_t]),
        "vfo_rle_decode": (I, [P, C.c_size_t, P, P, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _dims(shape):
    return np.asarray(shape, np.uint32)


def voxelize(xyz, tri, aabb_min, aabb_max, dims, parts=False):
    xyz = np.ascontiguousarray(xyz, np.float32)
    tri = np.ascontiguousarray(tri, np.uint32)
    d = _dims(dims)
    occ = np.zeros(tuple(dims), np.uint8)
    shell = np.zeros(tuple(dims), np.uint8) if parts else None
    par = np.zeros(tuple(dims), np.uint8) if parts else None
    n = lib().vfo_voxelize(_p(xyz), xyz.shape[0], _p(tri), tri.shape[0], _p(np.asarray(aabb_min, np.float32)),
                           _p(np.asarray(aabb_max, np.float32)), _p(d), _p(occ),
                           _p(shell) if parts else None, _p(par) if parts else None)
    if parts:
        return occ, shell, par
    return occ


def sat(v0, v1, v2, center, half):
    f = lambda a: np.ascontiguousarray(a, np.float32)
    return bool(lib().vfo_sat_triangle_box(_p(f(v0)), _p(f(v1)), _p(f(v2)), _p(f(center)), _p(f(half))))


def occ_to_grid(occ):
    occ = np.ascontiguousarray(occ, np.uint8)
    g = np.empty(occ.shape, np.uint16)
    lib().vfo_occ_to_grid(_p(occ), _p(g), occ.size)
    return g


def naive(grid, seeds, dist):
    g = np.ascontiguousarray(grid, np.uint16)
    s = np.ascontiguousarray(np.asarray(seeds, np.uint32).reshape(-1, 4))
    out = np.empty_like(g)
    lib().vfo_naive(_p(g), _p(out), _p(_dims(g.shape)), _p(s), s.shape[0], dist)
    return out


def flood(grid, seeds, neighbourhood, noise_seed=0, amp=0, dijkstra=False):
    g = np.ascontiguousarray(grid, np.uint16)
    s = np.ascontiguousarray(np.asarray(seeds, np.uint32).reshape(-1, 4))
    out = np.empty_like(g)
    fn = lib().vfo_flood_dijkstra if dijkstra else lib().vfo_flood
    rc = fn(_p(g), _p(out), _p(_dims(g.shape)), _p(s), s.shape[0], neighbourhood, noise_seed, amp)
    if rc != 0:
        raise ValueError("oracle flood rejected its arguments")
    return out


def remove_isolated(grid, seeds):
    g = np.ascontiguousarray(grid, np.uint16)
    s = np.ascontiguousarray(np.asarray(seeds, np.uint32).reshape(-1, 4))
    out = np.empty_like(g)
    lib().vfo_remove_isolated(_p(g), _p(out), _p(_dims(g.shape)), _p(s), s.shape[0])
    return out


def detect_boundaries(grid, size=1):
    g = np.ascontiguousarray(grid, np.uint16)
    out = np.empty_like(g)
    lib().vfo_detect_boundaries(_p(g), _p(out), _p(_dims(g.shape)), size)
    return out


def undo_mask(grid):
    g = np.array(grid, np.uint16, copy=True)
    lib().vfo_undo_mask(_p(g), g.size)
    return g


def homogenize(grid):
    g = np.array(grid, np.uint16, copy=True)
    lib().vfo_homogenize(_p(g), g.size)
    return g


def histogram(grid):
    g = np.ascontiguousarray(grid, np.uint16)
    out = np.zeros(65536, np.uint32)
    lib().vfo_histogram(_p(g), g.size, _p(out))
    return out


def noise_delta(seed, idx, amp):
    return lib().vfo_noise_delta(seed, idx, amp)


def hash32(key, counter):
    return lib().vfo_hash32(key, counter)


def seed_uniform(grid, n, seed, location=2):  # 2 = BOTH (Seeder::Location order)
    g = np.ascontiguousarray(grid, np.uint16)
    out = np.zeros((n, 4), np.uint32)
    rc = lib().vfo_seed_uniform(_p(g), _p(_dims(g.shape)), n, seed, location, _p(out))
    if rc != 0:
        raise ValueError(f"oracle seeding failed ({rc})")
    return out


def seed_merge(frags, seeds, dist):
    """Seeder::mergeSeeds: returns a copy of seeds with merged labels."""
    f = np.ascontiguousarray(np.asarray(frags, np.uint32).reshape(-1, 4))
    out = np.array(np.asarray(seeds, np.uint32).reshape(-1, 4), copy=True)
    rc = lib().vfo_seed_merge(_p(f), f.shape[0], _p(out), out.shape[0], dist)
    if rc != 0:
        raise ValueError(f"oracle merge failed ({rc})")
    return out


def seed_near(grid, frags, impacts, num_seeds, spreading, seed):
    """Seeder::nearSeeds (counter-based): fragments followed by the impact seeds."""
    g = np.ascontiguousarray(grid, np.uint16)
    f = np.ascontiguousarray(np.asarray(frags, np.uint32).reshape(-1, 4))
    out = np.zeros((f.shape[0] + num_seeds, 4), np.uint32)
    n = C.c_uint32()
    rc = lib().vfo_seed_near(_p(g), _p(_dims(g.shape)), _p(f), f.shape[0], impacts, num_seeds, spreading, seed,
                             _p(out), C.byref(n))
    if rc != 0:
        raise ValueError(f"oracle near seeding failed ({rc})")
    return out[: n.value]


def erosion_mask(kind, size):
    """(mask[side, side, side] uint8, activations float32)"""
    side = size if size % 2 else size + 1
    m = np.zeros(side ** 3, np.uint8)
    act = C.c_float()
    if lib().vfo_erosion_mask(kind, size, _p(m), C.byref(act)) < 0:
        raise ValueError("bad erosion type")
    return m.reshape(side, side, side), np.float32(act.value)


def erosion_noise(seed, cell):
    return np.float32(lib().vfo_erosion_noise(seed, cell))


def erode_step(grid, mask, limit, probability, noise_seed):
    g = np.ascontiguousarray(grid, np.uint16)
    m = np.ascontiguousarray(mask, np.uint8)
    out = np.empty_like(g)
    lib().vfo_erode_step(_p(g), _p(out), _p(_dims(g.shape)), _p(m), m.shape[0], float(limit), float(probability),
                         noise_seed)
    return out


def remove_isolated_cells(grid):
    g = np.ascontiguousarray(grid, np.uint16)
    out = np.empty_like(g)
    lib().vfo_remove_isolated_cells(_p(g), _p(out), _p(_dims(g.shape)))
    return out


def erode(grid, kind, size, iterations, probability, threshold, noise_seed):
    g = np.array(grid, np.uint16, copy=True)
    if lib().vfo_erode(_p(g), _p(_dims(g.shape)), kind, size, iterations, float(probability), float(threshold),
                       noise_seed) != 0:
        raise ValueError("bad erosion arguments")
    return g


def marching_cubes(grid, aabb_min, aabb_max, label=0xFFFFFFFF, smooth=None):
    """smooth: optional (nonboundary_iterations, nonboundary_weight, boundary_iterations, boundary_weight)."""
    g = np.ascontiguousarray(grid, np.uint16)
    m = Mesh()
    lib().vfo_marching_cubes(_p(g), _p(_dims(g.shape)), _p(np.asarray(aabb_min, np.float32)),
                             _p(np.asarray(aabb_max, np.float32)), label, C.byref(m))
    if smooth is not None:
        lib().vfo_mesh_smooth(C.byref(m), int(smooth[0]), float(smooth[1]), int(smooth[2]), float(smooth[3]))
    nv, nt, nl = m.nv, m.nt, m.nl

    def arr(ptr, n, dt):
        if n == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

    out = {
        "xyz": arr(m.xyz, 3 * nv, np.float32).reshape(nv, 3),
        "xyz64": arr(m.xyz64, 3 * nv, np.float64).reshape(nv, 3),
        "edge": arr(m.edge, nv, np.uint64),
        "tri": arr(m.tri, 3 * nt, np.uint32).reshape(nt, 3),
        "flag": arr(m.tri_flag, nt, np.uint8),
        "cube": arr(m.tri_cube, nt, np.uint64),
        "labels": arr(m.labels, nl, np.uint16),
        "voff": arr(m.voff, nl + 1, np.uint32),
        "toff": arr(m.toff, nl + 1, np.uint32),
    }
    lib().vfo_mesh_free(C.byref(m))
    return out


def mc_packed(config):
    return lib().vfo_mc_packed(config)


def rle_encode(grid):
    g = np.ascontiguousarray(grid, np.uint16)
    d = _dims(g.shape)
    n = lib().vfo_rle_encode(_p(g), _p(d), None, 0)
    buf = np.empty(n, np.uint8)
    lib().vfo_rle_encode(_p(g), _p(d), _p(buf), n)
    return buf.tobytes()


def rle_decode(data: bytes):
    buf = np.frombuffer(data, np.uint8)
    dims = np.zeros(3, np.uint32)
    hdr = np.frombuffer(data[:12], np.uint32)
    V = int(hdr[0]) * int(hdr[1]) * int(hdr[2])
    g = np.empty(V, np.uint16)
    rc = lib().vfo_rle_decode(_p(np.ascontiguousarray(buf)), buf.size, _p(dims), _p(g), V)
    if rc != 0:
        raise ValueError(f"RLE decode failed ({rc})")
    return g.reshape(tuple(int(x) for x in dims))


def checksum(grid):
    """Order-independent grid checksum, the same definition as the product's: the sum
    over non-EMPTY cells c holding v of (a << 32) | fmix32(a ^ c ^ 0x7F4A7C15) with
    a = fmix32(c ^ v * 0x9E3779B1)."""
    g = np.ascontiguousarray(grid, np.uint16).ravel()
    c = np.nonzero(g)[0].astype(np.uint32)
    v = g[c].astype(np.uint32)
    with np.errstate(over="ignore"):
        a = _fmix32(c ^ (v * np.uint32(0x9E3779B1)).astype(np.uint32))
        b = _fmix32(a ^ c ^ np.uint32(0x7F4A7C15))
    terms = (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64)
    return int(np.sum(terms, dtype=np.uint64))


def _fmix32(h):
    h = h.astype(np.uint32)
    h ^= h >> np.uint32(16)
    h = (h * np.uint32(0x85EBCA6B)).astype(np.uint32)
    h ^= h >> np.uint32(13)
    h = (h * np.uint32(0xC2B2AE35)).astype(np.uint32)
    h ^= h >> np.uint32(16)
    return h


def _hash32_np(key, counter):
    key = np.asarray(key, np.uint64)
    lo = (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    hi = (key >> np.uint64(32)).astype(np.uint32)
    with np.errstate(over="ignore"):
        a = _fmix32(np.asarray(counter, np.uint32) ^ lo)
        b = (a + (hi * np.uint32(0x9E3779B1)).astype(np.uint32) + np.uint32(0x7F4A7C15)).astype(np.uint32)
        return _fmix32(b)
