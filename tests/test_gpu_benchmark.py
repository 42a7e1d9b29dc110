"""The JSONL harness on the backend: the smoke matrix at 128x72 in latency
mode, every frame record's checksum equal to the SHA-256 of the oracle's
frame for the same source (bit-exact output through the harness), the
records parse with the published key sets, a second run compared against
the first matches its checksums, and throughput mode fills its window."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from lgcr_amd import benchmark as bm

pytestmark = pytest.mark.gpu

FIXTURE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "benchmark_schema.json")))
W, H = 128, 72


@pytest.fixture(autouse=True)
def restore_process_state(monkeypatch):
    # the harness sets LGCR_PROFILE and pins the process's CPU affinity
    monkeypatch.delenv("LGCR_PROFILE", raising=False)
    cpus = os.sched_getaffinity(0)
    yield
    os.sched_setaffinity(0, cpus)


def read(path):
    return [json.loads(line) for line in open(path) if line.strip()]


def oracle_digest(case, n):
    src = bm.Source(case, n + 1, static=False)
    y, u, v = src.planes(n)
    kw = case.kwargs()
    if case.operation == "Downsample":
        out = oracle.downsample_frame(oracle.downsample_params(kw["quality"], kw["kernel"]), y, u, v)
        planes = (y, out[1], out[2])
    else:
        p = oracle.make_params(W, H, kernel=kw["kernel"], taps=kw.get("taps"), strength=kw["strength"],
                               sparse=kw["sparse"], algo=kw.get("algo", 2))
        out = oracle.recon_frame(p, y, u, v, ss_w=case.ss_w, ss_h=case.ss_h)
        planes = (y, out[1], out[2])
    h = hashlib.sha256()
    for a in planes:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def test_latency_smoke_matrix_matches_oracle(device, tmp_path):
    first = tmp_path / "first.jsonl"
    assert bm.main(["--width", str(W), "--height", str(H), "--iterations", "2", "--output", str(first)]) == 0
    records = read(first)
    meta, frames = records[0], [r for r in records if r["record_type"] == "frame"]
    summaries = [r for r in records if r["record_type"] == "summary"]
    assert meta["record_type"] == "metadata" and meta["schema"] == 2
    cases = {c.key: c for c in bm.latency_cases("smoke", W, H)}
    assert [s["case"] for s in summaries] == list(cases)
    assert len(frames) == 2 * len(cases)
    for r in frames:
        for ks in FIXTURE["key_sets"]["frame"]:
            assert set(ks) <= set(r)
        case = cases[r["case"]]
        assert r["checksum"] == oracle_digest(case, r["frame"]), f"{r['case']} frame {r['frame']}"
        assert r["plugin_total_ms"] > 0 and r["wall_ms"] > 0
        if case.operation == "Recon":
            assert r["output_pixels"] == W * H
            if case.algo == "2":
                assert r["stage_us"]["reconstructChroma"] > 0 and r["taps_visited"] > 0
                assert 0.0 < r["sparse_active_ratio"] < 1.0
        else:
            assert r["stage_us"]["downsample_base"] > 0
    for s in summaries:
        for ks in FIXTURE["key_sets"]["summary/latency"]:
            assert set(ks) <= set(s)

    # a second run against the first: same outputs, speedups reported
    second = tmp_path / "second.jsonl"
    assert bm.main(["--width", str(W), "--height", str(H), "--iterations", "2", "--filter", "lanczos3/",
                    "--baseline", str(first), "--output", str(second)]) == 0
    again = [r for r in read(second) if r["record_type"] == "summary"]
    assert again and all(s["baseline_checksum_match"] for s in again)
    assert all(s["speedup"] > 0 for s in again)


def test_throughput_mode(device, tmp_path):
    out = tmp_path / "tp.jsonl"
    assert bm.main(["--mode", "throughput", "--width", str(W), "--height", str(H), "--threads", "2",
                    "--duration", "0.5", "--warmup", "0.2", "--output", str(out)]) == 0
    records = read(out)
    assert records[0]["mode"] == "throughput" and records[0]["inflight"] == 4
    summaries = records[1:]
    assert [s["case"] for s in summaries] == [f"{W}x{H}/420/lanczos3/{a}/sparse=1/" for a in "246"]
    for s in summaries:
        for ks in FIXTURE["key_sets"]["summary/throughput"]:
            assert set(ks) <= set(s)
        assert s["frames"] > 0 and s["fps"] > 0
        assert s["completion_latency_p50_ms"] <= s["completion_latency_p95_ms"]
        assert sum(s["stage_median_us"].values()) > 0
