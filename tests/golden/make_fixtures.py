"""Writes tests/golden/rle_sample.json from the reference's shipped RLE sample.

The reference ships docs/decompress/samples/AL_12B_grid_128r.rle (a 128x110x128
occupancy grid in RegularGrid::exportRLE format) and its decoded companion
AL_12B_grid_128r.npy (raw uint8 bytes written by decompress_grid.py after
flip(axis=1) and zero padding to 128^3).  The .rle is copied next to this
script as data; the 2 MiB .npy is pinned by its SHA-256 and value counts.
Run from the repository root: python tests/golden/make_fixtures.py
"""
import hashlib
import json
import shutil
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/docs/decompress/samples")
HERE = Path(__file__).resolve().parent

if __name__ == "__main__":
    shutil.copyfile(SRC / "AL_12B_grid_128r.rle", HERE / "AL_12B_grid_128r.rle")
    raw = (SRC / "AL_12B_grid_128r.npy").read_bytes()
    a = np.frombuffer(raw, np.uint8)
    rec = {
        "rle": "AL_12B_grid_128r.rle",
        "rle_sha256": hashlib.sha256((HERE / "AL_12B_grid_128r.rle").read_bytes()).hexdigest(),
        "decoded_npy_sha256": hashlib.sha256(raw).hexdigest(),
        "decoded_npy_bytes": len(raw),
        "decoded_value_counts": {str(int(v)): int(c) for v, c in zip(*np.unique(a, return_counts=True))},
        "dims": [128, 110, 128],
        "padded_dims": [128, 128, 128],
    }
    (HERE / "rle_sample.json").write_text(json.dumps(rec, indent=2) + "\n")
    print(json.dumps(rec, indent=2))
