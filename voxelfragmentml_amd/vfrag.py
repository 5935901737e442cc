"""ctypes binding of libvfrag's C ABI (include/vfrag.h) for Python callers.

This is plumbing for tests, the bench and scripting; the product is the C ABI
and the HIP kernels behind it.  Loading fails loudly when libvfrag.so has not
been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libvfrag.so"

EUCLIDEAN, MANHATTAN, CHEBYSHEV = 0, 1, 2
NAIVE, FLOOD = 0, 1
NEIGH_FROM_DISTANCE, NEIGH_VON_NEUMANN, NEIGH_MOORE = 0, 6, 26
SEED_INNER, SEED_OUTER, SEED_BOTH = 0, 1, 2  # Seeder::Location order
EROSION_SQUARE, EROSION_ELLIPSE, EROSION_CROSS = 0, 1, 2  # FractureParameters::ErosionType order
ALL_LABELS = 0xFFFFFFFF

ERRORS = {
    -1: "VF_ERR_INVALID_ARG", -2: "VF_ERR_OUT_OF_MEMORY", -3: "VF_ERR_DEVICE", -4: "VF_ERR_TOO_MANY_SEEDS",
    -5: "VF_ERR_BUDGET", -6: "VF_ERR_CANCELLED", -7: "VF_ERR_NOT_READY", -8: "VF_ERR_OVERFLOW",
    -9: "VF_ERR_SEEDING", -10: "VF_ERR_POISON",
}


class VfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


class MeshInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("num_labels", C.c_uint32)]


class FragmentMeta(C.Structure):
    _fields_ = [
        ("id", C.c_uint32), ("value", C.c_uint32), ("voxels", C.c_uint32), ("occupied_voxels", C.c_uint32),
        ("percentage", C.c_float), ("voxelization_size", C.c_uint32 * 3),
    ]


class MeshOut(C.Structure):
    _fields_ = [
        ("xyz", C.c_void_p), ("tri", C.c_void_p), ("tri_flag", C.c_void_p), ("labels", C.c_void_p),
        ("vert_offsets", C.c_void_p), ("tri_offsets", C.c_void_p),
        ("cap_vertices", C.c_uint32), ("cap_triangles", C.c_uint32), ("cap_labels", C.c_uint32),
        ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("num_labels", C.c_uint32),
        ("copied", C.c_int),
    ]


class Job(C.Structure):
    _fields_ = [
        ("xyz", C.c_void_p), ("tri", C.c_void_p), ("nv", C.c_uint32), ("nt", C.c_uint32),
        ("dims", C.c_uint32 * 3), ("aabb_min", C.c_float * 3), ("aabb_max", C.c_float * 3),
        ("num_seeds", C.c_uint32), ("seed", C.c_uint64), ("seed_location", C.c_int),
        ("algorithm", C.c_int), ("distance", C.c_int), ("neighbourhood", C.c_int),
        ("noise_seed", C.c_uint64), ("noise_amp", C.c_uint32), ("boundary_size", C.c_int),
        ("marching_cubes", C.c_int), ("smooth_iterations", C.c_uint32 * 2), ("smooth_weights", C.c_float * 2),
        ("mesh_out", C.c_void_p),
        ("extra_seeds", C.c_uint32), ("extra_seed", C.c_uint64), ("merge_distance", C.c_int),
        ("erode", C.c_int), ("erosion_convolution", C.c_int), ("erosion_size", C.c_uint32),
        ("erosion_iterations", C.c_uint32), ("erosion_probability", C.c_float), ("erosion_threshold", C.c_float),
        ("erosion_seed", C.c_uint64),
        ("seeds", C.c_void_p), ("impacts", C.c_uint32), ("impact_seeds", C.c_uint32),
        ("impact_spreading", C.c_uint32), ("impact_seed", C.c_uint64),
        ("mesh_on_device", C.c_int), ("poison", C.c_int),
        ("labels_out", C.c_void_p), ("histogram_out", C.c_void_p),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int), ("voxels", C.c_uint64), ("occupied", C.c_uint64), ("checksum", C.c_uint64),
        ("num_labels", C.c_uint32), ("mc_vertices", C.c_uint32), ("mc_triangles", C.c_uint32),
        ("flood_levels", C.c_uint32), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
        ("flood_bricks", C.c_uint32), ("flood_sweeps", C.c_uint32),
    ]


def lib():
    """Load libvfrag.so (built in-tree by voxelfragmentml_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m voxelfragmentml_amd.build`")
    L = C.CDLL(str(LIB_PATH))
    P, U32, U64, I, F = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float
    sig = {
        "vf_last_error": (C.c_char_p, []), "vf_version": (I, []), "vf_device_count": (I, []),
        "vf_grid_create": (I, [P, P, P, C.POINTER(P)]), "vf_grid_destroy": (I, [P]),
        "vf_grid_set_aabb": (I, [P, P, P, P]), "vf_grid_dims": (I, [P, P]),
        "vf_grid_read": (I, [P, P]), "vf_grid_write": (I, [P, P]), "vf_occupancy_read": (I, [P, P]),
        "vf_grid_count_occupied": (I, [P, C.POINTER(U64)]), "vf_grid_device_ptr": (I, [P, C.POINTER(P)]),
        "vf_grid_synchronize": (I, [P]),
        "vf_voxelize": (I, [P, P, U32, P, U32]), "vf_voxelize_device": (I, [P, P, U32, P, U32]),
        "vf_seed_uniform": (I, [P, U32, U64, I, P]),
        "vf_seed_near": (I, [P, P, U32, U32, U32, U32, U64, P, C.POINTER(U32)]),
        "vf_seed_merge": (I, [P, U32, P, U32, I]),
        "vf_fracture_naive": (I, [P, P, U32, I]), "vf_remove_isolated_regions": (I, [P, P, U32]),
        "vf_fracture_flood": (I, [P, P, U32, I, I, U64, U32]),
        "vf_detect_boundaries": (I, [P, I]), "vf_undo_mask": (I, [P]), "vf_homogenize": (I, [P]),
        "vf_erode": (I, [P, I, U32, U32, F, F, U64]), "vf_remove_isolated_cells": (I, [P]),
        "vf_histogram": (I, [P, P, U32]), "vf_fragment_metadata": (I, [P, P, P, U32, C.POINTER(U32)]),
        "vf_marching_cubes": (I, [P, U32, C.POINTER(MeshInfo)]),
        "vf_mesh_read": (I, [P, P, P, P, P, P, P]),
        "vf_mesh_smooth": (I, [P, U32, F, U32, F]),
        "vf_mesh_export_binm": (I, [P, U32, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "vf_grid_export_rle": (I, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "vf_grid_import_rle": (I, [P, P, C.c_size_t]),
        "vf_vessel_mesh": (I, [U64, U32, U32, P, C.POINTER(U32), P, C.POINTER(U32)]),
        "vf_batch_create": (I, [C.POINTER(P)]), "vf_batch_destroy": (I, [P]),
        "vf_batch_run": (I, [P, P, U32, P]), "vf_batch_set_graphs": (I, [P, I]),
        "vf_plain_run": (I, [P, U32, P]),
        "vf_server_create": (I, [C.c_size_t, U32, C.POINTER(P)]), "vf_server_destroy": (I, [P]),
        "vf_server_submit": (I, [P, P, C.POINTER(U64)]), "vf_server_cancel": (I, [P, U64]),
        "vf_server_poll": (I, [P, U64, P]), "vf_server_wait": (I, [P, U64, P]),
        "vf_server_stats": (I, [P, C.POINTER(C.c_size_t), C.POINTER(U32), C.POINTER(U64)]),
        "vf_server_memory": (I, [P] + [C.POINTER(C.c_size_t)] * 4),
        "vf_device_alloc": (I, [C.c_size_t, C.POINTER(P)]), "vf_device_free": (I, [P]),
        "vf_memcpy_htod": (I, [P, P, C.c_size_t]),
        "vf_timing_enable": (I, [I]), "vf_timing_reset": (I, []),
        "vf_timing_read": (I, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(U64)]),
        "vf_tune_run": (I, [C.c_char_p, I]), "vf_tune_load": (I, [C.c_char_p]),
        "vf_tune_jobs": (I, [P, U32, C.c_char_p, I, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int):
    if rc != 0:
        raise VfError(rc, lib().vf_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    return lib().vf_device_count()


def vessel_mesh(seed: int, radial: int, height_segments: int):
    """Closed lathe vessel: (xyz float32 (nv,3), tri uint32 (nt,3))."""
    L = lib()
    nv, nt = C.c_uint32(), C.c_uint32()
    check(L.vf_vessel_mesh(seed, radial, height_segments, None, C.byref(nv), None, C.byref(nt)))
    xyz = np.empty((nv.value, 3), np.float32)
    tri = np.empty((nt.value, 3), np.uint32)
    check(L.vf_vessel_mesh(seed, radial, height_segments, _ptr(xyz), C.byref(nv), _ptr(tri), C.byref(nt)))
    return xyz, tri


def vessel_segments(target_triangles: int):
    """(radial, height) segment counts giving about target_triangles (height = 2.5 x radial)."""
    r = max(3, int(round((target_triangles / 10.0) ** 0.5)))
    h = max(1, int(round(2.5 * r)))
    return r, h


def fragment_metadata(counts, dims) -> list:
    """FragmentationProcedure::FragmentMetadata entries (FragmentationProcedure.h:22-37) from the
    65536 label counts Grid.histogram returns: one dict per label above FREE with a nonzero count,
    in ascending label order (RegularGrid.cpp:443-471)."""
    c = np.ascontiguousarray(counts, np.uint32)
    if c.size != 65536:
        raise ValueError("counts must have 65536 entries")
    d = np.asarray(dims, np.uint32)
    n = C.c_uint32(0)
    check(lib().vf_fragment_metadata(_ptr(c), _ptr(d), None, 0, C.byref(n)))
    out = (FragmentMeta * max(1, n.value))()
    check(lib().vf_fragment_metadata(_ptr(c), _ptr(d), out, n.value, C.byref(n)))
    return [{"id": f.id, "value": f.value, "voxels": f.voxels, "occupied_voxels": f.occupied_voxels,
             "percentage": f.percentage, "voxelization_size": tuple(f.voxelization_size)}
            for f in out[:n.value]]


def seeds_array(seeds) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint32).reshape(-1, 4))
    return a


def seed_merge(frags, seeds, distance: int) -> np.ndarray:
    """Seeder::mergeSeeds: a copy of seeds labelled fragment | sub-seed id << 8."""
    f = seeds_array(frags)
    out = np.array(seeds_array(seeds), copy=True)
    check(lib().vf_seed_merge(_ptr(f), f.shape[0], _ptr(out), out.shape[0], distance))
    return out


class Grid:
    """RegularGrid on the device (MF/Source/DataStructures/RegularGrid.h)."""

    def __init__(self, dims, aabb_min, aabb_max):
        self.dims = tuple(int(d) for d in dims)
        self._h = C.c_void_p()
        d = np.asarray(self.dims, np.uint32)
        mn = np.asarray(aabb_min, np.float32)
        mx = np.asarray(aabb_max, np.float32)
        check(lib().vf_grid_create(_ptr(d), _ptr(mn), _ptr(mx), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().vf_grid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def V(self):
        return self.dims[0] * self.dims[1] * self.dims[2]

    def set_aabb(self, aabb_min, aabb_max, dims):
        d = np.asarray(dims, np.uint32)
        check(lib().vf_grid_set_aabb(self._h, _ptr(np.asarray(aabb_min, np.float32)),
                                     _ptr(np.asarray(aabb_max, np.float32)), _ptr(d)))
        self.dims = tuple(int(x) for x in dims)

    def read(self) -> np.ndarray:
        out = np.empty(self.dims, np.uint16)
        check(lib().vf_grid_read(self._h, _ptr(out)))
        return out

    def write(self, grid: np.ndarray):
        g = np.ascontiguousarray(grid, dtype=np.uint16)
        assert g.size == self.V
        check(lib().vf_grid_write(self._h, _ptr(g)))

    def occupancy(self) -> np.ndarray:
        out = np.empty(self.dims, np.uint8)
        check(lib().vf_occupancy_read(self._h, _ptr(out)))
        return out

    def count_occupied(self) -> int:
        n = C.c_uint64()
        check(lib().vf_grid_count_occupied(self._h, C.byref(n)))
        return n.value

    def voxelize(self, xyz: np.ndarray, tri: np.ndarray):
        xyz = np.ascontiguousarray(xyz, np.float32)
        tri = np.ascontiguousarray(tri, np.uint32)
        check(lib().vf_voxelize(self._h, _ptr(xyz), xyz.shape[0], _ptr(tri), tri.shape[0]))

    def voxelize_device(self, mesh: "DeviceMesh"):
        check(lib().vf_voxelize_device(self._h, mesh.d_xyz, mesh.nv, mesh.d_tri, mesh.nt))

    def seed_uniform(self, n: int, seed: int, location: int = SEED_BOTH) -> np.ndarray:
        out = np.zeros((n, 4), np.uint32)
        check(lib().vf_seed_uniform(self._h, n, seed, location, _ptr(out)))
        return out

    def seed_near(self, frags, impacts: int, num_seeds: int, spreading: int, seed: int) -> np.ndarray:
        """Seeder::nearSeeds: the fragment seeds followed by the impact seeds."""
        f = seeds_array(frags)
        out = np.zeros((f.shape[0] + num_seeds, 4), np.uint32)
        n = C.c_uint32()
        check(lib().vf_seed_near(self._h, _ptr(f), f.shape[0], impacts, num_seeds, spreading, seed, _ptr(out),
                                 C.byref(n)))
        return out[: n.value]

    def fracture_naive(self, seeds, distance: int):
        s = seeds_array(seeds)
        check(lib().vf_fracture_naive(self._h, _ptr(s), s.shape[0], distance))

    def remove_isolated_regions(self, seeds):
        s = seeds_array(seeds)
        check(lib().vf_remove_isolated_regions(self._h, _ptr(s), s.shape[0]))

    def fracture_flood(self, seeds, distance: int, neighbourhood: int = NEIGH_FROM_DISTANCE,
                       noise_seed: int = 0, noise_amp: int = 0):
        s = seeds_array(seeds)
        check(lib().vf_fracture_flood(self._h, _ptr(s), s.shape[0], distance, neighbourhood, noise_seed, noise_amp))

    def detect_boundaries(self, size: int = 1):
        check(lib().vf_detect_boundaries(self._h, size))

    def undo_mask(self):
        check(lib().vf_undo_mask(self._h))

    def erode(self, convolution: int, size: int, iterations: int, probability: float, threshold: float,
              noise_seed: int = 0):
        """RegularGrid::erode: boundary detection + erosion steps, then the isolated-cell cleanup."""
        check(lib().vf_erode(self._h, convolution, size, iterations, probability, threshold, noise_seed))

    def remove_isolated_cells(self):
        """RegularGrid::removeIsolatedRegions (the grid pass)."""
        check(lib().vf_remove_isolated_cells(self._h))

    def homogenize(self):
        check(lib().vf_homogenize(self._h))

    def histogram(self, n: int = 65536) -> np.ndarray:
        out = np.zeros(n, np.uint32)
        check(lib().vf_histogram(self._h, _ptr(out), n))
        return out

    def fragment_metadata(self) -> list:
        """Per-fragment records of RegularGrid::toTriangleMesh (RegularGrid.cpp:443-471)."""
        return fragment_metadata(self.histogram(), self.dims)

    def marching_cubes_info(self, label: int = ALL_LABELS) -> dict:
        """Marching cubes on the device; only the sizes come back (read_mesh copies the mesh out)."""
        info = MeshInfo()
        check(lib().vf_marching_cubes(self._h, label, C.byref(info)))
        self._mesh_sizes = (info.num_vertices, info.num_triangles, info.num_labels)
        return {"num_vertices": info.num_vertices, "num_triangles": info.num_triangles, "num_labels": info.num_labels}

    def marching_cubes(self, label: int = ALL_LABELS) -> dict:
        info = MeshInfo()
        check(lib().vf_marching_cubes(self._h, label, C.byref(info)))
        nv, nt, nl = info.num_vertices, info.num_triangles, info.num_labels
        self._mesh_sizes = (nv, nt, nl)
        xyz = np.empty((nv, 3), np.float32)
        tri = np.empty((nt, 3), np.uint32)
        flag = np.empty(nt, np.uint8)
        labels = np.empty(nl, np.uint16)
        voff = np.empty(nl + 1, np.uint32)
        toff = np.empty(nl + 1, np.uint32)
        check(lib().vf_mesh_read(self._h, _ptr(xyz), _ptr(tri), _ptr(flag), _ptr(labels), _ptr(voff), _ptr(toff)))
        return {"xyz": xyz, "tri": tri, "flag": flag, "labels": labels, "voff": voff, "toff": toff}

    def smooth_mesh(self, nonboundary_iterations: int, nonboundary_weight: float, boundary_iterations: int,
                    boundary_weight: float):
        """Laplacian smoothing of the last marching_cubes mesh, in place (read it again afterwards)."""
        check(lib().vf_mesh_smooth(self._h, nonboundary_iterations, nonboundary_weight, boundary_iterations,
                                   boundary_weight))

    def read_mesh(self) -> dict:
        """The current mesh of the last marching_cubes call (after any smoothing)."""
        info_nv, info_nt, nl = self._mesh_sizes
        xyz = np.empty((info_nv, 3), np.float32)
        tri = np.empty((info_nt, 3), np.uint32)
        flag = np.empty(info_nt, np.uint8)
        labels = np.empty(nl, np.uint16)
        voff = np.empty(nl + 1, np.uint32)
        toff = np.empty(nl + 1, np.uint32)
        check(lib().vf_mesh_read(self._h, _ptr(xyz), _ptr(tri), _ptr(flag), _ptr(labels), _ptr(voff), _ptr(toff)))
        return {"xyz": xyz, "tri": tri, "flag": flag, "labels": labels, "voff": voff, "toff": toff}

    def export_binm(self, label_index: int) -> bytes:
        """One fragment mesh of the last marching_cubes call in the reference's .binm layout."""
        size = C.c_size_t()
        check(lib().vf_mesh_export_binm(self._h, label_index, None, 0, C.byref(size)))
        buf = np.empty(size.value, np.uint8)
        check(lib().vf_mesh_export_binm(self._h, label_index, _ptr(buf), buf.size, C.byref(size)))
        return buf.tobytes()

    def export_rle(self) -> bytes:
        size = C.c_size_t()
        check(lib().vf_grid_export_rle(self._h, None, 0, C.byref(size)))
        buf = np.empty(size.value, np.uint8)
        check(lib().vf_grid_export_rle(self._h, _ptr(buf), buf.size, C.byref(size)))
        return buf.tobytes()

    def import_rle(self, data: bytes):
        buf = np.frombuffer(data, np.uint8)
        check(lib().vf_grid_import_rle(self._h, _ptr(buf), buf.size))


def timing_enable(on: bool = True):
    check(lib().vf_timing_enable(1 if on else 0))


def timing_reset():
    check(lib().vf_timing_reset())


def timing_read(kernel: str):
    ms, n = C.c_double(), C.c_uint64()
    check(lib().vf_timing_read(kernel.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value


def tune_run(path, quick: bool = True):
    """Run the autotuner on this device and write the launch-shape table to path."""
    check(lib().vf_tune_run(str(path).encode(), 1 if quick else 0))


def tune_load(path):
    """Install a launch-shape table written by tune_run."""
    check(lib().vf_tune_load(str(path).encode()))


def tune_jobs(jobs, path=None, quick: bool = True):
    """Tune the launch shapes of the bucket this batch uses on the batch itself; returns the median batch
    milliseconds (before, after) and writes the table to path when given."""
    arr = (Job * len(jobs))(*jobs)
    before, after = C.c_double(), C.c_double()
    check(lib().vf_tune_jobs(C.cast(arr, C.c_void_p), len(jobs), str(path).encode() if path else None,
                             1 if quick else 0, C.byref(before), C.byref(after)))
    return before.value, after.value


TUNING_FIELDS = ("max_edge", "cell_blocks", "flood_blocks", "pair_blocks", "brick_blocks", "naive_run", "brick_run",
                 "round_chunk", "lead_chunks", "mc_blocks")


def read_tuning(path):
    """The table of a tuning file: {bucket: {max_edge, cell_blocks, ..., lead_chunks}}."""
    out = {}
    for line in Path(path).read_text().splitlines():
        if line.startswith("#") or not line.strip():
            continue
        v = [int(x) for x in line.split()]
        out[v[0]] = dict(zip(TUNING_FIELDS, v[1:]))
    return out


def write_tuning(path, table):
    """Write a table in read_tuning's form (for vf_tune_load)."""
    lines = [" ".join(str(x) for x in [b] + [table[b][k] for k in TUNING_FIELDS]) for b in sorted(table)]
    Path(path).write_text("\n".join(lines) + "\n")


# ------------------------------------------------------------ batched / served
def make_job(xyz, tri, dims, aabb_min, aabb_max, num_seeds, seed, algorithm, distance,
             neighbourhood=NEIGH_FROM_DISTANCE, noise_seed=0, noise_amp=0, boundary_size=1,
             marching_cubes=0, seed_location=SEED_BOTH, poison=0, keep=None, smooth=None):
    """Fill a vf_job.  `keep` (a list) receives the arrays the job points to so
    they outlive the call.  smooth: optional (nonboundary_iterations,
    nonboundary_weight, boundary_iterations, boundary_weight) for the extracted
    meshes."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    tri = np.ascontiguousarray(tri, np.uint32)
    if keep is not None:
        keep.extend([xyz, tri])
    j = Job()
    j.xyz = xyz.ctypes.data
    j.tri = tri.ctypes.data
    j.nv, j.nt = xyz.shape[0], tri.shape[0]
    j.dims[:] = [int(d) for d in dims]
    j.aabb_min[:] = [float(v) for v in aabb_min]
    j.aabb_max[:] = [float(v) for v in aabb_max]
    j.num_seeds, j.seed, j.seed_location = num_seeds, seed, seed_location
    j.algorithm, j.distance, j.neighbourhood = algorithm, distance, neighbourhood
    j.noise_seed, j.noise_amp = noise_seed, noise_amp
    j.boundary_size, j.marching_cubes, j.poison = boundary_size, marching_cubes, poison
    if smooth is not None:
        j.smooth_iterations[:] = [int(smooth[0]), int(smooth[2])]
        j.smooth_weights[:] = [float(smooth[1]), float(smooth[3])]
    j.mesh_on_device = 0
    return j


class MeshBuffers:
    """Host buffers a job's marching-cubes mesh is copied into (vf_mesh_out), with numpy views."""

    def __init__(self, cap_vertices, cap_triangles, cap_labels=4096):
        self.xyz = np.zeros((cap_vertices, 3), np.float32)
        self.tri = np.zeros((cap_triangles, 3), np.uint32)
        self.flag = np.zeros(cap_triangles, np.uint8)
        self.labels = np.zeros(cap_labels, np.uint16)
        self.voff = np.zeros(cap_labels + 1, np.uint32)
        self.toff = np.zeros(cap_labels + 1, np.uint32)
        self.out = MeshOut()
        o = self.out
        o.xyz, o.tri, o.tri_flag = self.xyz.ctypes.data, self.tri.ctypes.data, self.flag.ctypes.data
        o.labels, o.vert_offsets, o.tri_offsets = self.labels.ctypes.data, self.voff.ctypes.data, self.toff.ctypes.data
        o.cap_vertices, o.cap_triangles, o.cap_labels = cap_vertices, cap_triangles, cap_labels

    def attach(self, job):
        job.mesh_out = C.addressof(self.out)
        return job

    def mesh(self):
        """The copied mesh in the layout Grid.marching_cubes returns (None when it was not copied)."""
        o = self.out
        if not o.copied:
            return None
        nv, nt, nl = o.num_vertices, o.num_triangles, o.num_labels
        return {"xyz": self.xyz[:nv].copy(), "tri": self.tri[:nt].copy(), "flag": self.flag[:nt].copy(),
                "labels": self.labels[:nl].copy(), "voff": self.voff[:nl + 1].copy(), "toff": self.toff[:nl + 1].copy()}


def _job_array(jobs):
    arr = (Job * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = j
    return arr


def result_dict(r: Result) -> dict:
    return {k: getattr(r, k) for k, _ in Result._fields_}


class Batch:
    """Batched form: every stage is one launch over all jobs (vf_batch_run)."""

    def __init__(self, graphs=True):
        self._h = C.c_void_p()
        check(lib().vf_batch_create(C.byref(self._h)))
        if not graphs:
            check(lib().vf_batch_set_graphs(self._h, 0))

    def run(self, jobs):
        arr = _job_array(jobs)
        res = (Result * len(jobs))()
        check(lib().vf_batch_run(self._h, arr, len(jobs), res))
        return [result_dict(r) for r in res]

    def close(self):
        if self._h:
            lib().vf_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plain_run(jobs):
    """Plain form: one job per launch (vf_plain_run)."""
    arr = _job_array(jobs)
    res = (Result * len(jobs))()
    check(lib().vf_plain_run(arr, len(jobs), res))
    return [result_dict(r) for r in res]


class Server:
    """Served form: continuous batching with HBM-budget admission."""

    def __init__(self, hbm_budget_bytes: int, max_slots: int = 64):
        self._h = C.c_void_p()
        check(lib().vf_server_create(hbm_budget_bytes, max_slots, C.byref(self._h)))

    def submit(self, job) -> int:
        t = C.c_uint64()
        check(lib().vf_server_submit(self._h, C.byref(job), C.byref(t)))
        return t.value

    def poll(self, ticket):
        r = Result()
        rc = lib().vf_server_poll(self._h, ticket, C.byref(r))
        if rc == -7:
            return None
        check(rc)
        return result_dict(r)

    def wait(self, ticket):
        r = Result()
        check(lib().vf_server_wait(self._h, ticket, C.byref(r)))
        return result_dict(r)

    def cancel(self, ticket):
        check(lib().vf_server_cancel(self._h, ticket))

    def stats(self):
        b, f, c = C.c_size_t(), C.c_uint32(), C.c_uint64()
        check(lib().vf_server_stats(self._h, C.byref(b), C.byref(f), C.byref(c)))
        return {"reserved_bytes": b.value, "in_flight": f.value, "completed": c.value}

    def memory(self):
        r, d, p, m = C.c_size_t(), C.c_size_t(), C.c_size_t(), C.c_size_t()
        check(lib().vf_server_memory(self._h, C.byref(r), C.byref(d), C.byref(p), C.byref(m)))
        return {"reserved_bytes": r.value, "device_bytes": d.value, "peak_device_bytes": p.value,
                "max_unreserved_bytes": m.value}

    def close(self):
        if self._h:
            lib().vf_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """A mesh copied once into HBM; jobs built from it set mesh_on_device."""

    def __init__(self, xyz, tri):
        self.xyz = np.ascontiguousarray(xyz, np.float32)
        self.tri = np.ascontiguousarray(tri, np.uint32)
        self.nv, self.nt = self.xyz.shape[0], self.tri.shape[0]
        self.d_xyz, self.d_tri = C.c_void_p(), C.c_void_p()
        check(lib().vf_device_alloc(self.xyz.nbytes, C.byref(self.d_xyz)))
        check(lib().vf_device_alloc(self.tri.nbytes, C.byref(self.d_tri)))
        check(lib().vf_memcpy_htod(self.d_xyz, _ptr(self.xyz), self.xyz.nbytes))
        check(lib().vf_memcpy_htod(self.d_tri, _ptr(self.tri), self.tri.nbytes))

    def job(self, *args, **kw):
        j = make_job(self.xyz, self.tri, *args, **kw)
        j.xyz, j.tri, j.mesh_on_device = self.d_xyz.value, self.d_tri.value, 1
        return j

    def close(self):
        for p in (self.d_xyz, self.d_tri):
            if p:
                lib().vf_device_free(p)
        self.d_xyz, self.d_tri = C.c_void_p(), C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
