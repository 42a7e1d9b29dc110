"""Latency and throughput harness over the plugin surface, writing the
reference's JSONL schema 2 so its records, tooling and published baselines
stay readable side by side with this backend's.

Records (one JSON object per line, keys sorted):

* ``metadata`` first (test/benchmark.py:530-549);
* latency mode: one ``frame`` record per measured frame (:624-654) and one
  ``summary`` per case (:656-696), with the ``baseline_*`` / ``speedup`` /
  ``stage_speedup`` / ``baseline_checksum_match`` comparison when
  ``--baseline`` names a prior JSONL file;
* throughput mode: one ``summary`` per case (:412-440) from a window of
  ``--inflight`` outstanding requests kept full for ``--duration`` seconds,
  with ``speedup`` and ``scaling_efficiency`` against the baseline (:559-569).

Cases and their keys are the reference's (``WxH/420/kernel/algo/sparse=N/``,
test/benchmark.py:84-89, matrices :161-237), so a reference baseline file
(``test/baselines/*.jsonl``) can be passed to ``--baseline`` directly.

The source follows the reference's synthetic scene (test/benchmark.py:130-
141): float 4:4:4 planes with a luma edge, a diagonal and a low-amplitude
texture, and a per-frame term of ``(n % 17) * 1e-7``.  4:2:0 and 4:2:2
sources are reduced here by a bilinear 2:1 filter (taps 1/4, 1/2, 1/4 on the
left-sited horizontal axis, 1/8, 3/8, 3/8, 1/8 on the centred vertical
axis, clamped at the borders); the reference reduces through zimg, so output
checksums are comparable between runs of this harness, not with the
reference's files.  Every source frame is prepared before its request is
timed (``source_precached`` is true in every record).  Frames go through
``lgcr_amd.plugin`` on host planes: the wall time includes the upload and
the download of each frame, as the reference's includes its conversions.
``scalar_max_error`` stays null: no scalar build is part of this backend.

Usage: ``python -m lgcr_amd.benchmark --preset standard --output run.jsonl``
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import gc
import hashlib
import json
import os
import platform
import queue
import subprocess
import sys
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import abi, paths, plugin

SCHEMA = 2

# stage names of the records' stage maps (stageName, src/pipeline.h:222-244)
STAGES = tuple(plugin.PROP_STAGE_NAMES)
CPU_SLOTS = tuple(plugin.CPU_SLOTS)

# kernel label -> Recon arguments beyond kernel=label
KERNEL_ARGS = {
    "bilinear": {}, "bicubic": {}, "spline16": {}, "spline36": {},
    "lanczos3": {"kernel": "lanczos", "taps": 3},
    "lanczos4": {"kernel": "lanczos", "taps": 4},
    "jinc3": {"kernel": "jinc", "taps": 3},
}
SUBSAMPLINGS = ((1, 1, "420"), (1, 0, "422"), (0, 0, "444"))
DOWNSAMPLE_KERNELS = ("spline36", "lanczos3", "binomial")


@dataclass(frozen=True)
class Case:
    width: int
    height: int
    ss_w: int
    ss_h: int
    subsampling: str
    kernel: str
    algo: str                 # "plain", "2", "4", "6", or "q<quality>" for Downsample
    sparse: bool
    options: tuple = ()
    operation: str = "Recon"

    @property
    def key(self):
        op = "" if self.operation == "Recon" else self.operation + "/"
        extra = ",".join(f"{k}={v}" for k, v in self.options)
        return f"{self.width}x{self.height}/{self.subsampling}/{op}{self.kernel}/{self.algo}/" \
               f"sparse={int(self.sparse)}/{extra}"

    def kwargs(self):
        if self.operation == "Downsample":
            return dict(self.options)
        kw = {"kernel": self.kernel}
        kw.update(KERNEL_ARGS[self.kernel])
        kw["strength"] = 0.0 if self.algo == "plain" else 0.8
        kw["sparse"] = int(self.sparse)
        if self.algo != "plain":
            kw["algo"] = int(self.algo)
        kw.update(self.options)
        return kw


def _sizes(preset, width, height):
    if width or height:
        return [(width or 1920, height or 1080)]
    return {"full": [(1920, 1080), (3840, 2160)], "standard": [(1920, 1080)]}.get(preset, [(640, 360)])


def _downsample_case(w, h, kernel, quality):
    return Case(w, h, 0, 0, "444", kernel, f"q{quality}", False,
                (("quality", quality), ("kernel", kernel)), "Downsample")


def latency_cases(preset="smoke", width=None, height=None):
    """The latency matrix (test/benchmark.py:161-219): smoke runs every
    kernel's plain base, algos 2/4/6 sparse and Downsample at three
    qualities; standard and full cover every format x kernel x algo x sparse
    plus the control ablations."""
    sizes = _sizes(preset, width, height)
    out = []
    if preset == "smoke":
        w, h = sizes[0]
        out += [Case(w, h, 1, 1, "420", k, "plain", False) for k in KERNEL_ARGS]
        out += [Case(w, h, 1, 1, "420", "lanczos3", a, True) for a in ("2", "4", "6")]
        out += [_downsample_case(w, h, "spline36", q) for q in (0, 1, 2)]
        return out
    for w, h in sizes:
        for ssw, ssh, label in SUBSAMPLINGS:
            for k in KERNEL_ARGS:
                out.append(Case(w, h, ssw, ssh, label, k, "plain", False))
                out += [Case(w, h, ssw, ssh, label, k, a, s) for a in ("2", "4", "6") for s in (False, True)]
        for name, values in (("ms", (0.0, 1.0)), ("ridge", (0, 1)), ("cedge", (0, 1)), ("bm", (0, 1))):
            out += [Case(w, h, 1, 1, "420", "lanczos3", "2", True, ((name, v),)) for v in values]
        out += [Case(w, h, 1, 1, "420", "lanczos3", "6", True, (("qgate", v),)) for v in (0.0, 1.0)]
        out += [_downsample_case(w, h, k, q) for k in DOWNSAMPLE_KERNELS for q in (0, 1, 2)]
    return out


def throughput_cases(preset="smoke", width=None, height=None):
    """The scaling matrix (test/benchmark.py:222-233): 4:2:0 lanczos3,
    algos 2/4/6 sparse."""
    return [Case(w, h, 1, 1, "420", "lanczos3", a, True)
            for w, h in _sizes(preset, width, height) for a in ("2", "4", "6")]


# ---------------------------------------------------------------------------
# source
# ---------------------------------------------------------------------------

def scene_planes(width, height):
    """Float 4:4:4 (Y, U, V) of the benchmark scene (test/benchmark.py:130-141)."""
    yy, xx = np.mgrid[0:height, 0:width]
    edge = (xx > width // 2).astype(np.float32)
    diagonal = (xx * height > yy * width).astype(np.float32)
    texture = (0.01 * np.sin(xx * 0.071) * np.cos(yy * 0.053)).astype(np.float32)
    return ((0.2 + 0.25 * edge + 0.03 * diagonal + texture).astype(np.float32),
            (-0.25 + 0.55 * edge + 0.5 * texture).astype(np.float32),
            (0.30 - 0.60 * edge - 0.5 * texture).astype(np.float32))


def _reduce_axis(a, axis, centred):
    """2:1 bilinear reduction along one axis in float64, borders clamped:
    left-sited outputs sit on even samples (taps 1/4, 1/2, 1/4), centred ones
    between samples 2j and 2j+1 (taps 1/8, 3/8, 3/8, 1/8)."""
    n = a.shape[axis]
    m = n // 2

    def take(offsets):
        idx = np.clip(2 * np.arange(m)[:, None] + np.asarray(offsets)[None, :], 0, n - 1)
        return np.take(a, idx, axis=axis)      # the new tap axis follows `axis`

    if centred:
        taps, w = take([-1, 0, 1, 2]), np.array([0.125, 0.375, 0.375, 0.125])
    else:
        taps, w = take([-1, 0, 1]), np.array([0.25, 0.5, 0.25])
    shape = [1] * taps.ndim
    shape[axis + 1] = w.size
    return (taps * w.reshape(shape)).sum(axis=axis + 1)


def reduce_chroma(plane, ss_w, ss_h):
    a = plane.astype(np.float64)
    if ss_w:
        a = _reduce_axis(a, 1, centred=False)
    if ss_h:
        a = _reduce_axis(a, 0, centred=True)
    return np.ascontiguousarray(a.astype(np.float32))


class Source:
    """Float source clip of a case: frame n is the scene plus (n % 17) * 1e-7,
    its chroma reduced to the case's subsampling.  ``static`` serves frame
    0's planes for every request (the reference's cached, looped source)."""

    def __init__(self, case: Case, length: int, static: bool):
        self.case = case
        self.base = scene_planes(case.width, case.height)
        self.fmt = plugin.query_video_format(plugin.YUV, plugin.FLOAT, 32, case.ss_w, case.ss_h)
        self.static = static
        self.length = length
        self._cache = {}
        self._lock = threading.Lock()

    def node(self):
        """The source clip (holds this Source; the Source does not hold it)."""
        return plugin.VideoNode(self.fmt, self.case.width, self.case.height, self.length, self.frame)

    def planes(self, n):
        key = 0 if self.static else n
        with self._lock:
            if key not in self._cache:
                d = np.float32((key % 17) * 1e-7)
                y, u, v = (p + d for p in self.base)
                self._cache[key] = (np.ascontiguousarray(y), reduce_chroma(u, self.case.ss_w, self.case.ss_h),
                                    reduce_chroma(v, self.case.ss_w, self.case.ss_h))
            return self._cache[key]

    def precache(self, n):
        self.planes(n)

    def forget(self, n):
        if not self.static:
            with self._lock:
                self._cache.pop(n, None)

    def frame(self, n):
        return plugin.VideoFrame(self.fmt, list(self.planes(n)), {})


def make_node(case: Case, source: Source):
    fn = getattr(plugin.core.lgcr, case.operation)
    return fn(source.node(), **case.kwargs())


# ---------------------------------------------------------------------------
# per-frame observations
# ---------------------------------------------------------------------------

def checksum(frame):
    """First 16 hex digits of SHA-256 over the three planes' bytes."""
    h = hashlib.sha256()
    for p in range(3):
        h.update(np.ascontiguousarray(frame[p]).tobytes())
    return h.hexdigest()[:16]


def frame_profile(frame):
    """(stage us, stage pixels, stage taps, cpu slot us) from _LGCR_* props."""
    pr = frame.props
    us = {s: float(pr.get(f"_LGCR_{s}_us", 0.0)) for s in STAGES}
    px = {s: int(pr.get(f"_LGCR_{s}_pixels", 0)) for s in STAGES}
    taps = {s: int(pr.get(f"_LGCR_{s}_taps", 0)) for s in STAGES}
    cpu = {s: float(pr.get(f"_LGCR_cpu_{s}_us", 0.0)) for s in CPU_SLOTS}
    return us, px, taps, cpu


def status_kib(field):
    try:
        with open("/proc/self/status", encoding="ascii") as f:
            for line in f:
                if line.startswith(field + ":"):
                    return int(line.split()[1])
    except OSError:
        pass
    return None


def reset_peak_rss():
    try:
        with open("/proc/self/clear_refs", "w", encoding="ascii") as f:
            f.write("5\n")
    except OSError:
        pass


def percentile(values, q):
    return float(np.percentile(values, q)) if values else None


def _median_map(samples, names):
    return {n: (float(np.median(samples[n])) if samples[n] else 0.0) for n in names}


# ---------------------------------------------------------------------------
# the two modes
# ---------------------------------------------------------------------------

def run_latency_case(case: Case, iterations: int, static: bool, profile: bool, emit):
    """Frames 2 .. iterations+1 one at a time after two warm-up frames;
    emits a frame record per frame, returns the summary."""
    source = Source(case, iterations + 2, static)
    node = make_node(case, source)
    try:
        for n in (0, 1):
            source.precache(n)
            node.get_frame(n)
            source.forget(n)
        reset_peak_rss()
        walls, totals, checksums = [], [], []
        stage_s = {s: [] for s in STAGES}
        cpu_s = {s: [] for s in CPU_SLOTS}
        kwargs = case.kwargs()
        for n in range(2, iterations + 2):
            source.precache(n)
            rss0 = status_kib("VmRSS")
            t0 = time.perf_counter_ns()
            frame = node.get_frame(n)
            wall_ms = (time.perf_counter_ns() - t0) / 1e6
            rss1 = status_kib("VmRSS")
            source.forget(n)
            us, px, taps, cpu = frame_profile(frame)
            total_ms = float(frame.props.get("_LGCR_total_us", 0.0)) / 1000.0
            digest = checksum(frame)
            walls.append(wall_ms)
            totals.append(total_ms)
            checksums.append(digest)
            for s in STAGES:
                stage_s[s].append(us[s])
            for s in CPU_SLOTS:
                cpu_s[s].append(cpu[s])
            emit({
                "record_type": "frame", "case": case.key, "frame": n, "config": kwargs,
                "operation": case.operation, "source_precached": True,
                "width": case.width, "height": case.height, "subsampling": case.subsampling,
                "kernel": case.kernel, "algo": case.algo, "sparse": case.sparse,
                "wall_ms": wall_ms, "plugin_total_ms": total_ms,
                "stage_us": us, "stage_pixels": px, "stage_taps": taps, "cpu_profile_us": cpu,
                "checksum": digest, "rss_kib": rss1,
                "rss_delta_kib": rss1 - rss0 if rss0 is not None and rss1 is not None else None,
                "peak_rss_kib": status_kib("VmHWM"),
                "output_pixels": int(frame.props.get("_LGCR_output_pixels", 0)),
                "taps_visited": int(frame.props.get("_LGCR_taps_visited", 0)),
                "sparse_active_ratio": float(frame.props.get("_LGCR_sparse_active_ratio", 1.0)),
                "scalar_max_error": None,
            })
        return {
            "record_type": "summary", "case": case.key, "operation": case.operation,
            "source_precached": True, "frames": iterations,
            "median_ms": float(np.median(walls)), "min_ms": min(walls),
            "plugin_total_median_ms": float(np.median(totals)) if profile else None,
            "stage_median_us": _median_map(stage_s, STAGES),
            "cpu_profile_median_us": _median_map(cpu_s, CPU_SLOTS),
            "checksums": checksums, "scalar_max_error": None, "peak_rss_kib": status_kib("VmHWM"),
        }
    finally:
        _close(node)


def _window(node, pool, start, inflight, duration, collect):
    """Keep `inflight` requests outstanding for `duration` seconds; returns
    (samples of the requests that completed inside the window, next frame
    number, window seconds).  Completion times are taken when a request
    finishes, not when this thread gets to it; requests still running at the
    deadline are drained without refilling."""
    finished = queue.SimpleQueue()            # (future, submit ns, completion ns)
    active = {}
    nxt = start
    samples = []

    def submit():
        nonlocal nxt
        t = time.perf_counter_ns()
        fut = pool.submit(node.get_frame, nxt)
        active[fut] = nxt
        nxt += 1
        fut.add_done_callback(lambda f, t=t: finished.put((f, t, time.perf_counter_ns())))

    def take(item, deadline):
        fut, t, c = item
        n = active.pop(fut)
        frame = fut.result()
        if c > deadline:
            return
        s = {"frame": n, "latency_ms": (c - t) / 1e6}
        if collect:
            us, _, _, cpu = frame_profile(frame)
            s.update(stage_us=us, cpu_profile_us=cpu,
                     plugin_total_ms=float(frame.props.get("_LGCR_total_us", 0.0)) / 1000.0)
        samples.append(s)

    begin = time.perf_counter_ns()
    deadline = begin + int(duration * 1e9)
    for _ in range(inflight):
        submit()
    while True:
        left = (deadline - time.perf_counter_ns()) / 1e9
        if left <= 0:
            break
        try:
            item = finished.get(timeout=left)
        except queue.Empty:
            break
        take(item, deadline)
        if time.perf_counter_ns() < deadline:
            submit()
    while active:
        take(finished.get(), deadline)
    return samples, nxt, (deadline - begin) / 1e9


def run_throughput_case(case: Case, threads: int, inflight: int, duration: float, warmup: float, profile: bool):
    """Completion-latency percentiles and frames/s over a window of `inflight`
    outstanding requests served by `threads` workers (test/benchmark.py:391-440)."""
    source = Source(case, 1 << 40, static=True)
    source.precache(0)
    node = make_node(case, source)
    try:
        with cf.ThreadPoolExecutor(max_workers=threads) as pool:
            _, nxt, _ = _window(node, pool, 0, inflight, warmup, False)
            reset_peak_rss()
            samples, _, elapsed = _window(node, pool, nxt, inflight, duration, profile)
    finally:
        _close(node)
    if not samples:
        raise RuntimeError(f"no frames completed for {case.key}")
    lat = [s["latency_ms"] for s in samples]
    stage_s = {s: [x["stage_us"][s] for x in samples if "stage_us" in x] for s in STAGES}
    cpu_s = {s: [x["cpu_profile_us"][s] for x in samples if "cpu_profile_us" in x] for s in CPU_SLOTS}
    totals = [x["plugin_total_ms"] for x in samples if "plugin_total_ms" in x]
    return {
        "record_type": "summary", "case": case.key, "operation": case.operation,
        "source_precached": True, "mode": "throughput", "threads": threads, "inflight": inflight,
        "duration_s": elapsed, "frames": len(samples), "fps": len(samples) / elapsed,
        "median_ms": percentile(lat, 50), "min_ms": min(lat),
        "completion_latency_p50_ms": percentile(lat, 50),
        "completion_latency_p95_ms": percentile(lat, 95),
        "plugin_total_median_ms": percentile(totals, 50),
        "stage_median_us": {s: (percentile(v, 50) if v else 0.0) for s, v in stage_s.items()},
        "cpu_profile_median_us": {s: (percentile(v, 50) if v else 0.0) for s, v in cpu_s.items()},
        "checksums": [], "scalar_max_error": None, "peak_rss_kib": status_kib("VmHWM"),
    }


def _close(node):
    b = getattr(node, "_backend", None)
    if b is not None:
        b.close()


# ---------------------------------------------------------------------------
# baseline comparison
# ---------------------------------------------------------------------------

def load_baseline(path):
    """Summary records of a prior JSONL file, by case key."""
    out = {}
    if path:
        with open(path, encoding="utf-8") as f:
            for line in f:
                line = line.strip()
                if line:
                    r = json.loads(line)
                    if r.get("record_type") == "summary":
                        out[r["case"]] = r
    return out


def compare_latency(summary, prior):
    """Adds the latency-mode comparison fields (test/benchmark.py:670-692)."""
    if not prior:
        return summary
    old = float(prior["median_ms"])
    summary["baseline_ms"] = old
    summary["speedup"] = old / summary["median_ms"]
    ps = prior.get("stage_median_us", {})
    summary["stage_baseline_us"] = ps
    mine = summary["stage_median_us"]
    summary["stage_speedup"] = {s: float(v) / mine[s] for s, v in ps.items()
                                if s in mine and float(v) > 0.0 and mine[s] > 0.0}
    pc = prior.get("checksums", [])
    if pc:
        k = min(len(pc), len(summary["checksums"]))
        summary["baseline_checksum_match"] = pc[:k] == summary["checksums"][:k]
    if prior.get("peak_rss_kib") is not None:
        summary["baseline_peak_rss_kib"] = prior["peak_rss_kib"]
    return summary


def compare_throughput(summary, prior, threads):
    """Adds the throughput-mode comparison fields (test/benchmark.py:559-569)."""
    if prior and prior.get("fps"):
        ref_threads = int(prior.get("threads", 1))
        sp = summary["fps"] / float(prior["fps"])
        summary.update(baseline_fps=float(prior["fps"]), baseline_threads=ref_threads, speedup=sp,
                       scaling_efficiency=sp / (threads / ref_threads))
    return summary


# ---------------------------------------------------------------------------
# command line
# ---------------------------------------------------------------------------

def parse_cpu_list(text):
    cpus = set()
    for item in (t.strip() for t in text.split(",")):
        if not item:
            continue
        if "-" in item:
            a, b = (int(x) for x in item.split("-", 1))
            if a < 0 or b < a:
                raise ValueError(f"invalid CPU range {item!r}")
            cpus.update(range(a, b + 1))
        else:
            c = int(item)
            if c < 0:
                raise ValueError(f"invalid CPU {c}")
            cpus.add(c)
    if not cpus:
        raise ValueError("CPU list is empty")
    return sorted(cpus)


def _first_line(cmd):
    try:
        return subprocess.check_output(cmd, text=True, stderr=subprocess.STDOUT).splitlines()[0]
    except (OSError, subprocess.CalledProcessError, IndexError):
        return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo", encoding="ascii") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or None


def metadata(args, cpus):
    throughput = args.mode == "throughput"
    return {
        "record_type": "metadata", "schema": SCHEMA, "mode": args.mode, "preset": args.preset,
        "iterations": args.iterations, "platform": platform.platform(),
        "python": platform.python_version(), "vapoursynth": None,
        "profile": not args.no_profile,
        "cpu": cpus[0] if len(cpus) == 1 else None, "cpu_list": cpus,
        "threads": args.threads, "inflight": args.inflight if throughput else None,
        "duration_s": args.duration if throughput else None,
        "warmup_s": args.warmup if throughput else None,
        "cpu_model": _cpu_model(),
        "compiler": _first_line([os.environ.get("HIPCC", "/opt/rocm/bin/hipcc"), "--version"]),
        "build_config": f"hip=gfx950;abi={abi.lib().lgcr_abi_version()};library={paths.HIP_LIB}",
        "backend": "liblgcr_hip",
    }


def build_parser():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0],
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--mode", choices=("latency", "throughput"), default="latency")
    ap.add_argument("--preset", choices=("smoke", "standard", "full"), default="smoke")
    ap.add_argument("--width", type=int)
    ap.add_argument("--height", type=int)
    ap.add_argument("--iterations", type=int, default=9)
    ap.add_argument("--cpu", type=int, help="logical CPU to pin to (default: first available)")
    ap.add_argument("--cpu-list", help="CPU affinity list/ranges, for example 0-7,16")
    ap.add_argument("--threads", type=int, default=1, help="worker threads issuing frame requests")
    ap.add_argument("--inflight", type=int, help="requests kept in flight (default: 2 x threads)")
    ap.add_argument("--duration", type=float, default=10.0, help="throughput measurement duration in seconds")
    ap.add_argument("--warmup", type=float, default=2.0, help="throughput warm-up duration in seconds")
    ap.add_argument("--source-mode", choices=("auto", "dynamic", "static"), default="auto")
    ap.add_argument("--namespace", default="lgcr", help="plugin namespace (only lgcr exists here)")
    ap.add_argument("--output", default="-")
    ap.add_argument("--baseline", help="prior JSONL file to compare")
    ap.add_argument("--compare-scalar", action="store_true",
                    help="not available: no scalar build is part of this backend")
    ap.add_argument("--no-profile", action="store_true",
                    help="measure production wall time without LGCR instrumentation")
    ap.add_argument("--filter", default="", help="substring matched against case id")
    ap.add_argument("--list", action="store_true")
    return ap


def validate(ap, args):
    """The reference harness's argument checks (test/benchmark.py:478-517)."""
    if args.threads < 1:
        ap.error("--threads must be positive")
    if args.inflight is None:
        args.inflight = 2 * args.threads
    if args.inflight < 1:
        ap.error("--inflight must be positive")
    if args.duration <= 0.0:
        ap.error("--duration must be positive")
    if args.warmup <= 0.0:
        ap.error("--warmup must be positive")
    if args.iterations < 1:
        ap.error("--iterations must be positive")
    if args.cpu is not None and args.cpu_list:
        ap.error("--cpu and --cpu-list are mutually exclusive")
    if args.compare_scalar:
        ap.error("--compare-scalar: no scalar build is part of this backend")
    if args.namespace != "lgcr":
        ap.error("--namespace: only the lgcr namespace exists here")
    if args.mode == "throughput" and args.source_mode == "dynamic":
        ap.error("throughput mode requires --source-mode auto or static")


def select_cpus(ap, args):
    have = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else []
    try:
        cpus = (parse_cpu_list(args.cpu_list) if args.cpu_list else
                [args.cpu] if args.cpu is not None else have[:args.threads])
    except ValueError as e:
        ap.error(str(e))
    if cpus and hasattr(os, "sched_setaffinity"):
        missing = sorted(set(cpus) - set(have)) if have else []
        if missing:
            ap.error(f"CPUs {missing} are outside available affinity {have}")
        if len(cpus) < args.threads:
            ap.error("CPU list must contain at least --threads CPUs")
        os.sched_setaffinity(0, set(cpus))
    return cpus


def main(argv=None):
    ap = build_parser()
    args = ap.parse_args(argv)
    validate(ap, args)
    matrix = (throughput_cases if args.mode == "throughput" else latency_cases)(args.preset, args.width,
                                                                                 args.height)
    cases = [c for c in matrix if args.filter in c.key]
    if args.list:
        for c in cases:
            print(c.key)
        return 0
    if not cases:
        ap.error("no benchmark cases matched")
    cpus = select_cpus(ap, args)
    if args.no_profile:
        os.environ.pop("LGCR_PROFILE", None)
    else:
        os.environ["LGCR_PROFILE"] = "1"
    baseline = load_baseline(args.baseline)
    out = sys.stdout if args.output == "-" else open(args.output, "w", encoding="utf-8")

    def emit(record):
        print(json.dumps(record, sort_keys=True), file=out, flush=True)

    try:
        emit(metadata(args, cpus))
        for case in cases:
            if args.mode == "throughput":
                s = run_throughput_case(case, args.threads, args.inflight, args.duration, args.warmup,
                                        not args.no_profile)
                emit(compare_throughput(s, baseline.get(case.key), args.threads))
            else:
                static = args.source_mode == "static"
                s = run_latency_case(case, args.iterations, static, not args.no_profile, emit)
                emit(compare_latency(s, baseline.get(case.key)))
            gc.collect()             # release the case's planes before the next case's peak-RSS window
    finally:
        if out is not sys.stdout:
            out.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
