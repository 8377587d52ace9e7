"""The label-keyed result store of the serving loop (host logic only):
skip-by-label as the reference's sweep resume does
(scripts/run_pending_sweep.py:207-240), atomic rewrites, duplicate labels
refused."""
import json
import os

import pytest

from lbm.serve import ResultStore


def test_store_round_trip_and_skip(tmp_path):
    path = str(tmp_path / "res.json")
    s = ResultStore(path)
    assert not s.done("a")
    s.add({"label": "a", "Dx": 3.0})
    s.add({"label": "b", "Dx": 4.0})
    assert not os.path.exists(path + ".tmp")
    t = ResultStore(path)
    assert t.done("a") and t.done("b") and not t.done("c")
    assert [r["label"] for r in t.results] == ["a", "b"]
    with pytest.raises(ValueError):
        t.add({"label": "a"})
    with open(path) as f:
        assert json.load(f) == t.results


def test_store_reads_reference_style_file(tmp_path):
    """A results file written by the reference's sweep driver (a JSON list of
    case dicts with a label) is read as-is: its labels count as done."""
    path = str(tmp_path / "thesis_sweep_results.json")
    with open(path, "w") as f:
        json.dump([{"label": "s3a_flat_We2.0", "max_k": 1.0},
                   {"label": "s5_ridge_R0.5_We5.0", "max_k": 1.2}], f)
    s = ResultStore(path)
    assert s.done("s3a_flat_We2.0") and s.done("s5_ridge_R0.5_We5.0")
    s.add({"label": "new", "max_k": 1.1})
    assert len(ResultStore(path).results) == 3
