"""bench.py's host-side pieces, without a GPU: the workload draw is the
stated distribution, the stage byte definitions add up to the path's, and the
JSON-lines job records are one valid JSON object per line."""
from __future__ import annotations

import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def test_job_specs_follow_the_mixed_distribution():
    specs = bench.job_specs(64, 4)
    assert len(specs) == 64
    for s in specs:
        assert 64 <= s["edge"] <= 384 and s["edge"] % 4 == 0
        assert 5000 <= s["tris"] <= 300000 and 2 <= s["seeds"] <= 64
        assert s["dist"] in (0, 1, 2) and s["algo"] in (0, 1)
    assert bench.job_specs(64, 4) == specs     # the same draw every run


def test_stage_bytes_sum_to_the_path_bytes():
    specs = bench.job_specs(16, 7)
    sb = bench.stage_bytes(specs)
    assert sum(sb.values()) == bench.path_bytes(specs)
    V = sum(s["edge"] ** 3 for s in specs)
    assert bench.stage_bytes(specs, 1000)["marching_cubes"] == 2 * V + 12000


def test_records_are_json_lines(tmp_path):
    assert not bench.Records("")
    p = tmp_path / "r.jsonl"
    r = bench.Records(str(p))
    specs = bench.job_specs(3, 1)
    for i, s in enumerate(specs):
        r.emit({"form": "batched", "job": i, **bench.spec_record(s), "stage_ms_alone": {"voxelize": 0.5}})
    r.emit({"form": "summary", "gvoxels_per_s": 1.0})
    r.close()
    lines = [json.loads(x) for x in p.read_text().splitlines()]
    assert [x["form"] for x in lines] == ["batched"] * 3 + ["summary"]
    for x, s in zip(lines, specs):
        assert x["dims"] == [s["edge"]] * 3 and x["seeds"] == s["seeds"]
        assert x["algorithm"] == ("naive" if s["algo"] == 0 else "flood")
        assert x["distance"] == ["euclidean", "manhattan", "chebyshev"][s["dist"]]
