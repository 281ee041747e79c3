"""CPU checks of bench.py's contract: one GPU only, and the HBM-traffic lookup, which reports
the bytes per update measured with rocprofv3 only while the library sources match the digest
they were measured on (a stale figure becomes null, never a wrong number)."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line


def test_traffic_lookup_follows_the_source_digest(tmp_path, monkeypatch):
    doc = {"sources_sha256": bench.sources_digest(), "bytes_per_update": {"lad": 1234.5}}
    path = tmp_path / "traffic.json"
    path.write_text(json.dumps(doc))
    monkeypatch.setattr(bench, "TRAFFIC_FILE", str(path))
    assert bench.measured_traffic("lad") == 1234.5
    assert bench.measured_traffic("plain") is None
    doc["sources_sha256"] = "0" * 64
    path.write_text(json.dumps(doc))
    assert bench.measured_traffic("lad") is None
    path.write_text("not json")
    assert bench.measured_traffic("lad") is None


def test_committed_traffic_file_is_well_formed():
    with open(bench.TRAFFIC_FILE) as f:
        doc = json.load(f)
    assert len(doc["sources_sha256"]) == 64
    per = doc["bytes_per_update"]
    assert set(per) >= {"plain", "lad", "learned"}
    # at least the step's algorithmic 640 B per update, and the shock-capturing step above the plain one
    assert 640.0 <= per["plain"] < per["lad"] < 4000.0
