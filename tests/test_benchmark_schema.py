"""The JSONL harness (lgcr_amd.benchmark) against the reference's published
schema-2 files: case keys, record key sets and profile-map names from
tests/golden/benchmark_schema.json (written from the reference's
test/baselines and test/results JSONL by make_benchmark_schema.py), the
baseline comparison arithmetic, the source's chroma siting and the command
line.  The harness runs here over a stand-in node, so no device is needed;
tests/test_gpu_benchmark.py runs it on the backend."""
import json
import os
import time

import numpy as np
import pytest

from lgcr_amd import benchmark as bm
from lgcr_amd import plugin

FIXTURE = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "benchmark_schema.json")))


def sizes_in_fixture():
    return sorted({tuple(int(v) for v in c.split("/")[0].split("x")) for c in FIXTURE["cases"]})


def test_every_published_case_key_is_in_a_matrix():
    ours = set()
    for w, h in sizes_in_fixture():
        for preset in ("smoke", "standard", "full"):
            ours.update(c.key for c in bm.latency_cases(preset, w, h))
            ours.update(c.key for c in bm.throughput_cases(preset, w, h))
    missing = [c for c in FIXTURE["cases"] if c not in ours]
    assert not missing, missing


def test_preset_matrices():
    smoke = bm.latency_cases("smoke")
    assert [c.key for c in smoke[:10]] == [c for c in FIXTURE["cases"] if c.startswith("640x360")][:10]
    assert len(smoke) == 13 and smoke[-1].operation == "Downsample"
    std = bm.latency_cases("standard")
    # 3 formats x 7 kernels x (plain + 3 algos x 2 sparse) + 8 ablations + 2 qgate + 9 Downsample
    assert len(std) == 3 * 7 * 7 + 8 + 2 + 9
    assert len(bm.latency_cases("full")) == 2 * len(std)
    assert [c.key for c in bm.throughput_cases("standard")] == [
        f"1920x1080/420/lanczos3/{a}/sparse=1/" for a in "246"]


def test_case_arguments():
    c = bm.Case(64, 32, 1, 1, "420", "lanczos4", "6", True, (("qgate", 0.0),))
    assert c.key == "64x32/420/lanczos4/6/sparse=1/qgate=0.0"
    assert c.kwargs() == {"kernel": "lanczos", "taps": 4, "strength": 0.8, "sparse": 1, "algo": 6, "qgate": 0.0}
    assert bm.Case(64, 32, 1, 1, "420", "bicubic", "plain", False).kwargs() == {
        "kernel": "bicubic", "strength": 0.0, "sparse": 0}
    d = bm._downsample_case(64, 32, "binomial", 2)
    assert d.key == "64x32/444/Downsample/binomial/q2/sparse=0/quality=2,kernel=binomial"
    assert d.kwargs() == {"quality": 2, "kernel": "binomial"}


def test_stage_and_slot_names_cover_the_published_maps():
    assert set(FIXTURE["stages"]) <= set(bm.STAGES)
    assert set(FIXTURE["cpu_slots"]) <= set(bm.CPU_SLOTS)
    # the newest published files carry every stage and slot
    assert set(FIXTURE["stages"]) == set(bm.STAGES)
    assert set(FIXTURE["cpu_slots"]) == set(bm.CPU_SLOTS)


# ---------------------------------------------------------------------------
# source
# ---------------------------------------------------------------------------

def test_chroma_reduction_siting():
    h, w = 12, 16
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    # a ramp along x: left-sited outputs sit on the even samples
    r = bm.reduce_chroma(xx.astype(np.float32), 1, 0)
    assert r.shape == (h, w // 2)
    np.testing.assert_allclose(r[:, 1:], xx[:, 2:w:2], atol=1e-6)
    assert r[0, 0] == pytest.approx(0.25)          # clamped left border: (0 + 2*0 + 1) / 4
    # a ramp along y: centred outputs sit half-way between rows 2j and 2j+1
    r = bm.reduce_chroma(yy.astype(np.float32), 0, 1)
    assert r.shape == (h // 2, w)
    np.testing.assert_allclose(r[1:-1], yy[2:-2:2] + 0.5, atol=1e-6)
    # both axes, constants kept, dtype float32
    r = bm.reduce_chroma(np.full((h, w), 0.3, np.float32), 1, 1)
    assert r.dtype == np.float32 and r.shape == (h // 2, w // 2)
    np.testing.assert_allclose(r, 0.3, atol=1e-7)
    # 4:4:4 is untouched
    a = np.random.default_rng(0).random((h, w), dtype=np.float32)
    np.testing.assert_array_equal(bm.reduce_chroma(a, 0, 0), a)


def test_source_frames():
    case = bm.Case(40, 24, 1, 1, "420", "lanczos3", "2", True)
    src = bm.Source(case, 20, static=False)
    f3, f20 = src.frame(3), src.frame(19)
    assert [p.shape for p in f3.planes] == [(24, 40), (12, 20), (12, 20)]
    assert all(p.dtype == np.float32 for p in f3.planes)
    y0 = bm.scene_planes(40, 24)[0]
    np.testing.assert_array_equal(f3[0], y0 + np.float32(3e-7))
    np.testing.assert_array_equal(f20[0], y0 + np.float32(2e-7))      # 19 % 17 = 2
    # the scene: luma edge at the centre column, chroma of opposite signs
    assert f3[1][0, 0] < 0 < f3[1][0, -1] and f3[2][0, 0] > 0 > f3[2][0, -1]
    static = bm.Source(case, 20, static=True)
    np.testing.assert_array_equal(static.frame(7)[0], static.frame(0)[0])


# ---------------------------------------------------------------------------
# the two modes over a stand-in node
# ---------------------------------------------------------------------------

class StandInNode:
    """Recon-shaped node: 4:4:4 float output with _LGCR_* props; a frame
    takes `delay` seconds."""

    def __init__(self, source, delay=0.0):
        self.source, self.delay, self.calls = source, delay, 0

    def get_frame(self, n):
        self.calls += 1
        if self.delay:
            time.sleep(self.delay)
        f = self.source.frame(n)
        out = [f[0], np.zeros_like(f[0]), np.full_like(f[0], n)]
        props = {f"_LGCR_{s}_us": 10.0 + i for i, s in enumerate(bm.STAGES)}
        props.update({f"_LGCR_{s}_pixels": 100 for s in bm.STAGES})
        props.update({f"_LGCR_{s}_taps": 7 for s in bm.STAGES})
        props.update({f"_LGCR_cpu_{s}_us": 0.0 for s in bm.CPU_SLOTS})
        props.update(_LGCR_total_us=1500.0, _LGCR_output_pixels=f[0].size, _LGCR_taps_visited=42,
                     _LGCR_sparse_active_ratio=0.25)
        return plugin.VideoFrame(f.format, out, props)


@pytest.fixture
def stand_in(monkeypatch):
    nodes = []

    def make(case, source):
        nodes.append(StandInNode(source, delay=0.002))
        return nodes[-1]

    monkeypatch.setattr(bm, "make_node", make)
    return nodes


def keys_cover(record, published):
    """Every key of every published version of this record kind is present."""
    for ks in published:
        missing = set(ks) - set(record)
        assert not missing, missing


def test_latency_records(stand_in):
    case = bm.Case(32, 16, 1, 1, "420", "lanczos3", "2", True)
    records = []
    summary = bm.run_latency_case(case, 3, False, True, records.append)
    assert [r["frame"] for r in records] == [2, 3, 4]
    assert stand_in[0].calls == 5                       # two warm-up frames, then three timed
    for r in records:
        keys_cover(r, FIXTURE["key_sets"]["frame"])
        assert r["case"] == case.key and r["config"] == case.kwargs()
        assert r["stage_us"]["buildGuideMaps"] == 10.0 + bm.STAGES.index("buildGuideMaps")
        assert r["stage_taps"]["plainChroma"] == 7 and r["taps_visited"] == 42
        assert r["plugin_total_ms"] == 1.5 and r["sparse_active_ratio"] == 0.25
        assert len(r["checksum"]) == 16 and r["wall_ms"] > 0
    keys_cover(summary, FIXTURE["key_sets"]["summary/latency"])
    assert summary["checksums"] == [r["checksum"] for r in records]
    assert len(set(summary["checksums"])) == 3          # the frames differ (frame term, plane 2)
    assert summary["median_ms"] == pytest.approx(float(np.median([r["wall_ms"] for r in records])))
    json.dumps(summary)
    json.dumps(records)


def test_throughput_summary(stand_in):
    case = bm.Case(32, 16, 1, 1, "420", "lanczos3", "4", True)
    s = bm.run_throughput_case(case, threads=2, inflight=4, duration=0.3, warmup=0.05, profile=True)
    keys_cover(s, FIXTURE["key_sets"]["summary/throughput"])
    assert s["mode"] == "throughput" and s["threads"] == 2 and s["inflight"] == 4
    assert s["duration_s"] == pytest.approx(0.3, abs=1e-3)
    # two workers at 2 ms a frame: at most ~1000 frames/s, and most of that
    assert 300 < s["fps"] <= 1100
    assert s["completion_latency_p50_ms"] <= s["completion_latency_p95_ms"]
    assert s["min_ms"] >= 2.0                          # no request completes faster than a frame takes
    assert s["stage_median_us"]["backProject"] == 10.0 + bm.STAGES.index("backProject")
    assert s["plugin_total_median_ms"] == 1.5


def test_metadata_record():
    args = bm.build_parser().parse_args(["--mode", "throughput", "--threads", "2"])
    bm.validate(bm.build_parser(), args)
    m = bm.metadata(args, [0, 1])
    for ks in FIXTURE["key_sets"]["metadata"]:
        if "cpu_list" in ks:                           # the current harness's metadata
            keys_cover(m, [ks])
    assert m["schema"] == 2 and m["inflight"] == 4 and m["cpu"] is None and m["cpu_list"] == [0, 1]


# ---------------------------------------------------------------------------
# baseline comparison against the published summaries
# ---------------------------------------------------------------------------

def published(case, mode):
    for s in FIXTURE["summaries"]:
        if s["case"] == case and (s["fps"] is not None) == (mode == "throughput"):
            return s
    raise KeyError(case)


def test_latency_comparison_with_published_baseline(tmp_path):
    prior = dict(published("1920x1080/420/lanczos3/2/sparse=1/", "latency"), record_type="summary",
                 checksums=["a", "b"], peak_rss_kib=1000)
    path = tmp_path / "base.jsonl"
    path.write_text(json.dumps({"record_type": "metadata", "schema": 2}) + "\n" + json.dumps(prior) + "\n")
    base = bm.load_baseline(str(path))
    assert list(base) == [prior["case"]]
    mine = {"median_ms": 2.0, "stage_median_us": {s: 0.0 for s in bm.STAGES}, "checksums": ["a", "c", "d"]}
    mine["stage_median_us"]["buildGuideMaps"] = 20.0
    out = bm.compare_latency(mine, base[prior["case"]])
    assert out["baseline_ms"] == prior["median_ms"] == pytest.approx(89.767513)
    assert out["speedup"] == pytest.approx(prior["median_ms"] / 2.0)
    assert out["stage_speedup"] == {"buildGuideMaps": pytest.approx(prior["stage_median_us"]["buildGuideMaps"]
                                                                    / 20.0)}
    assert out["baseline_checksum_match"] is False and out["baseline_peak_rss_kib"] == 1000
    assert bm.compare_latency({"median_ms": 1.0}, None) == {"median_ms": 1.0}


def test_throughput_comparison_with_published_baseline():
    prior = published("1920x1080/420/lanczos3/2/sparse=1/", "throughput")
    t16 = [s for s in FIXTURE["summaries"] if s["case"] == prior["case"] and s["threads"] == 16][0]
    assert t16["fps"] == pytest.approx(260.6, abs=0.05)
    out = bm.compare_throughput({"fps": 10000.0}, t16, threads=8)
    assert out["baseline_fps"] == t16["fps"] and out["baseline_threads"] == 16
    assert out["speedup"] == pytest.approx(10000.0 / t16["fps"])
    assert out["scaling_efficiency"] == pytest.approx(out["speedup"] / 0.5)


# ---------------------------------------------------------------------------
# command line
# ---------------------------------------------------------------------------

def test_list_and_filter(capsys):
    assert bm.main(["--list", "--filter", "lanczos3/6"]) == 0
    assert capsys.readouterr().out.split() == ["640x360/420/lanczos3/6/sparse=1/"]
    assert bm.main(["--list", "--mode", "throughput", "--preset", "full"]) == 0
    assert len(capsys.readouterr().out.split()) == 6


@pytest.mark.parametrize("argv,needle", [
    (["--threads", "0"], "--threads must be positive"),
    (["--inflight", "0"], "--inflight must be positive"),
    (["--duration", "0"], "--duration must be positive"),
    (["--warmup", "-1"], "--warmup must be positive"),
    (["--iterations", "0"], "--iterations must be positive"),
    (["--cpu", "0", "--cpu-list", "0"], "mutually exclusive"),
    (["--compare-scalar"], "no scalar build"),
    (["--filter", "no-such-case"], "no benchmark cases matched"),
    (["--mode", "throughput", "--source-mode", "dynamic"], "throughput mode requires"),
])
def test_argument_errors(argv, needle, capsys):
    with pytest.raises(SystemExit):
        bm.main(argv)
    assert needle in capsys.readouterr().err


def test_cpu_list_parsing():
    assert bm.parse_cpu_list("0-3,8, 10") == [0, 1, 2, 3, 8, 10]
    for bad in ("", "3-1", "-2"):
        with pytest.raises(ValueError):
            bm.parse_cpu_list(bad)
