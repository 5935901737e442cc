"""The C ABI and the C++ drop-in surface, without a GPU: libvfrag.so loads and
exports every function include/*.h declares, the C++ mirror header compiles
with and without the reference's global names, and the drop-in caller is
built and answers a bad command line without touching the device."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
INCLUDE = ROOT / "include"
LIB = ROOT / "voxelfragmentml_amd" / "libvfrag.so"
DRIVER = ROOT / "voxelfragmentml_amd" / "dropin" / "fracture_scene"


def declared_functions():
    names = set()
    for h in INCLUDE.glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(vf_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_the_path():
    names = declared_functions()
    for must in ["vf_grid_create", "vf_voxelize", "vf_seed_uniform", "vf_fracture_naive", "vf_fracture_flood",
                 "vf_detect_boundaries", "vf_histogram", "vf_marching_cubes", "vf_batch_run", "vf_plain_run",
                 "vf_server_submit", "vf_tune_run", "vf_tune_load"]:
        assert must in names
    assert len(names) >= 45


def test_library_exports_every_declared_function():
    if not LIB.exists():
        pytest.skip("libvfrag.so not built")
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


@pytest.mark.parametrize("ref_names", [False, True])
def test_cpp_mirror_header_compiles(ref_names, tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(("#define VFRAG_REFERENCE_NAMES\n" if ref_names else "") + """
#include "vfrag_fracture.hpp"
int f(vfrag::RegularGrid& g, const std::vector<vfrag::uvec4>& seeds) {
    vfrag::FractureParameters p;
    vfrag::fracturer::Fracturer* fr = vfrag::fracturer::FloodFracturer::getInstance();
    if (!fr->setDistanceFunction(vfrag::fracturer::CHEBYSHEV_DISTANCE)) return 1;
    fr->build(g, seeds, &p);
    return (int)g.getNumSubdivisions().x;
}
""" + ("int h(RegularGrid& g) { FractureParameters p; return fracturer::NaiveFracturer::getInstance()->setDistanceFunction(fracturer::MANHATTAN_DISTANCE) ? (int)g.length() : 0; }\n" if ref_names else ""))
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{INCLUDE}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_distance_and_location_enums_match_reference():
    h = (INCLUDE / "vfrag.h").read_text()
    # FractureParameters::DistanceFunction {EUCLIDEAN, MANHATTAN, CHEBYSHEV}; Seeder::Location {INNER, OUTER, BOTH}
    for name, v in [("VF_EUCLIDEAN", 0), ("VF_MANHATTAN", 1), ("VF_CHEBYSHEV", 2),
                    ("VF_SEED_INNER", 0), ("VF_SEED_OUTER", 1), ("VF_SEED_BOTH", 2), ("VF_NAIVE", 0), ("VF_FLOOD", 1)]:
        assert re.search(rf"#define {name} {v}\b", h), name


def test_dropin_driver_usage_without_device():
    if not DRIVER.exists():
        pytest.skip("drop-in driver not built")
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.parametrize("dist", [0, 1, 2])
def test_seed_merge_host_entry_matches_oracle(dist):
    """vf_seed_merge is host code (the reference merges on the CPU too), so it
    runs here without a device."""
    if not LIB.exists():
        pytest.skip("libvfrag.so not built")
    import numpy as np
    import oracle as O
    from voxelfragmentml_amd import vfrag

    rng = np.random.default_rng(40 + dist)
    frags = np.concatenate([rng.integers(0, 256, (10, 3)), (2 + np.arange(10))[:, None]], 1).astype(np.uint32)
    extra = np.concatenate([rng.integers(0, 256, (20, 3)), np.full((20, 1), 2)], 1).astype(np.uint32)
    seeds = np.concatenate([frags, extra])
    np.testing.assert_array_equal(vfrag.seed_merge(frags, seeds, dist), O.seed_merge(frags, seeds, dist))
    many = np.zeros((256, 4), np.uint32)
    with pytest.raises(vfrag.VfError) as e:
        vfrag.seed_merge(frags[:1], many, dist)
    assert e.value.code == -4


def test_fragment_metadata_host_entry():
    """vf_fragment_metadata is host code over the label counts: the entries of
    RegularGrid::toTriangleMesh (RegularGrid.cpp:443-471), one per label above
    FREE, percentage = voxels / occupied voxels in float32."""
    if not LIB.exists():
        pytest.skip("libvfrag.so not built")
    import numpy as np
    from voxelfragmentml_amd import vfrag

    counts = np.zeros(65536, np.uint32)
    counts[[0, 1]] = [1000, 77]          # EMPTY and FREE: never part of the metadata
    counts[2], counts[3], counts[9] = 3, 1, 4
    meta = vfrag.fragment_metadata(counts, (10, 20, 30))
    assert [m["value"] for m in meta] == [2, 3, 9] and [m["id"] for m in meta] == [0, 1, 2]
    assert [m["voxels"] for m in meta] == [3, 1, 4]
    assert all(m["occupied_voxels"] == 8 and m["voxelization_size"] == (10, 20, 30) for m in meta)
    assert [m["percentage"] for m in meta] == [0.375, 0.125, 0.5]   # exact in binary

    rng = np.random.default_rng(9)
    labels = rng.integers(0, 40, 200000).astype(np.uint16)
    counts = np.bincount(labels, minlength=65536).astype(np.uint32)
    counts[[0, 1, 17]] = 0
    meta = vfrag.fragment_metadata(counts, (64, 64, 64))
    keep = np.flatnonzero(counts[2:]) + 2
    total = int(counts[2:].sum())
    assert [m["value"] for m in meta] == keep.tolist()
    expect = (counts[keep].astype(np.float32) / np.float32(total)).tolist()
    assert [m["percentage"] for m in meta] == expect
    assert abs(sum(m["percentage"] for m in meta) - 1.0) < 1e-5
    n = ctypes.c_uint32(0)
    out = (vfrag.FragmentMeta * 4)()
    dims = np.array([64, 64, 64], np.uint32)
    assert vfrag.lib().vf_fragment_metadata(vfrag._ptr(counts), vfrag._ptr(dims), out, 4, ctypes.byref(n)) == 0
    assert n.value == len(keep) and [out[i].value for i in range(4)] == keep[:4].tolist()
