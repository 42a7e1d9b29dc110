"""Benchmark: Recon 1920x1080 YUV420P10 -> YUV444P10, algo 2, bp 0.5.

One step = one pass of the Recon hot path over a batch of synthetic frames
that are already resident in HBM (default 240 frames, one scene per frame).
Prints ONE JSON line (rank 0).  The CPU baseline is the oracle (scalar C
restatement of the reference) timed on a bounded sample of the same
workload on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vapoursynth-lgcr_amd"))
sys.path.insert(0, ROOT)

METRIC = "Recon 1080p420 10-bit fps, HBM GB/s (% peak); serving p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0


VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9   # non-packed FP32 / FP64 VALU lane-ops per s


# Algorithmic VALU lane-ops of the tiled guided pass (plain base + guided
# correction), counted from the reference's scalar arithmetic: one op per
# float/double add, mul, fma, min/max, compare, conversion, sqrt or divide.
#   per guided tap: range/hull pass 6 (recon.cpp:296-330), weight and
#     accumulation 27 (recon.cpp:584-620)
#   per active output pixel: guide metadata 266 (four bilinear guide reads,
#     tensor direction, normal minimum, ridge probe, phase and mutual gates;
#     recon.cpp:155-215, lgcr.h:467-538) and adaptive sigma, fades, rescue
#     with kappa, normalisation and hull clamp 110 (recon.cpp:330-480, 713-724)
#   per separable plain-base tap and plane: 2 (recon.cpp:1144-1147, 1453-1459)
OPS_PER_GUIDED_TAP = 33
OPS_PER_GUIDED_PIXEL = 376
OPS_PER_PLAIN_TAP_PLANE = 2


def algorithmic_valu(clip, frames):
    """Mean algorithmic lane-ops per frame of the guided tiles over `frames`
    (host planes), from the backend's per-frame tap and workset counts."""
    from lgcr_amd import abi
    total = 0
    for f in frames:
        _, st = clip.recon_host(*f, stats=True)
        total += (OPS_PER_PLAIN_TAP_PLANE * 2 * st.stage_taps[abi.STAGE_BUILD_BASE_CHROMA]
                  + OPS_PER_GUIDED_TAP * st.stage_taps[abi.STAGE_APPLY_GUIDED_CORRECTION]
                  + OPS_PER_GUIDED_PIXEL * st.active_output_pixels)
    return total / max(1, len(frames))


def profile_figure(name, key):
    """`key = N` from a committed rocprofv3 summary under profiles/, or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key):
                return int(line.split("=")[1])
    except OSError:
        pass
    return None


def measured_traffic(workload_key):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same workload, or None."""
    return profile_figure(f"{workload_key}_hbm_traffic.md", "traffic_bytes_per_launch")


def measured_valu(workload_key):
    """VALU lane-ops (SQ_INSTS_VALU x 64) per launch of the dominant kernel
    from the committed rocprofv3 counter summary, or None."""
    return profile_figure(f"{workload_key}_valu.md", "valu_lane_ops_per_launch")


def host_path(clip, frames, n, threads):
    """Frames/s through host planes (upload, Recon, download per frame;
    pageable numpy buffers as the plugin hands them over), `threads` frames
    in flight."""
    import threading
    nxt = [0]
    lock = threading.Lock()

    def work():
        while True:
            with lock:
                k = nxt[0]
                nxt[0] += 1
            if k >= n:
                return
            clip.recon_host(*frames[k % len(frames)])

    ts = [threading.Thread(target=work) for _ in range(threads)]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return n / (time.perf_counter() - t0)


def algorithmic_bytes(w, h, cw, ch, bps):
    """Compulsory plane I/O per frame: Y, U, V read once; U, V 4:4:4 written
    once; same-size Recon shares Y (not written)."""
    return bps * (w * h + 2 * cw * ch + 2 * w * h)


def cpu_baseline(frames, params_kw, bits, seconds, threads):
    """Oracle frames/s on `threads` host threads for about `seconds`."""
    import threading
    import oracle
    w = frames[0][0].shape[1]
    h = frames[0][0].shape[0]
    done = [0] * threads
    stop = time.perf_counter() + seconds

    def work(t):
        p = oracle.make_params(w, h, **params_kw)
        k = t
        while time.perf_counter() < stop:
            y, u, v = frames[k % len(frames)]
            oracle.recon_frame(p, y, u, v, bits=bits)
            done[t] += 1
            k += threads

    t0 = time.perf_counter()
    ts = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    return sum(done) / el, sum(done), el


def downsample(args):
    """Downsample 1920x1080 YUV444P10 -> YUV420P10 (spline36, left siting,
    strength 1) at each quality on device-resident frames, one frame at a
    time (each call waits for its frame); quality 0 through host planes from
    eight threads; the oracle on one host thread."""
    import ctypes as C

    import numpy as np
    import torch
    from lgcr_amd import abi, synth
    W, H, n = 1920, 1080, 32
    frames = synth.clip(n, W, H, ss_w=0, ss_h=0, bits=10, seed=1, scene_per_frame=True,
                        threads=min(16, os.cpu_count() or 1))
    dev = abi.Device(0)
    din = [[torch.from_numpy(p.view(np.int16)).to("cuda") for p in f] for f in frames]
    dout = [torch.empty((H // 2, W // 2), dtype=torch.int16, device="cuda") for _ in range(2)]
    torch.cuda.synchronize()
    pin = [[(t.data_ptr(), t.stride(0) * 2) for t in f] for f in din]
    pout = [(0, 0)] + [(t.data_ptr(), t.stride(0) * 2) for t in dout]
    out = {"workload": "Downsample 1920x1080 YUV444P10 -> YUV420P10, spline36, left, strength 1",
           "frames": n, "inflight": 1}
    # algorithmic bytes: Y, U, V 4:4:4 read once, U, V 4:2:0 written once
    # (luma is shared, not written)
    nbytes = 3 * W * H * 2 + 2 * (W // 2) * (H // 2) * 2
    out["bytes_per_frame"] = nbytes
    L = abi.lib()
    for q in (0, 1, 2):
        clip = abi.DownsampleClip(dev, abi.downsample_params(quality=q), abi.make_format(np.uint16, W, H, 0, 0, 10))
        for f in pin[:4]:
            clip.downsample_device(f, pout)
        t0 = time.perf_counter()
        for f in pin:
            clip.downsample_device(f, pout)
        el = time.perf_counter() - t0
        out[f"q{q}_fps"] = round(n / el, 1)
        # per-stage device time (HIP events around each stage), separate pass
        abi.check(L.lgcr_clip_set_timing(clip.handle, 1))
        for f in pin:
            clip.downsample_device(f, pout)
        us = (C.c_double * abi.STAGE_COUNT)()
        cnt = (C.c_uint64 * abi.STAGE_COUNT)()
        abi.check(L.lgcr_clip_stage_times(clip.handle, us, cnt))
        stages = {abi.STAGE_NAMES[i]: round(us[i] / cnt[i], 2) for i in range(abi.STAGE_COUNT) if cnt[i]}
        dev_us = sum(stages.values())
        out[f"q{q}_stage_avg_us"] = stages
        out[f"q{q}_device_us"] = round(dev_us, 2)
        out[f"q{q}_hbm_frac_of_device_time"] = round(nbytes / (dev_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 5)
        clip.close()
    if not args.no_host_path:
        # host planes in and out per frame (the plugin boundary,
        # PCIe-inclusive), quality 0, eight frames in flight from eight host
        # threads; luma shared
        from concurrent.futures import ThreadPoolExecutor
        clip = abi.DownsampleClip(dev, abi.downsample_params(quality=0),
                                  abi.make_format(np.uint16, W, H, 0, 0, 10), max_inflight=8)

        def one(f):
            return clip.downsample_host(*f, copy_luma=False)

        with ThreadPoolExecutor(8) as ex:
            list(ex.map(one, frames[:8]))
            t0 = time.perf_counter()
            list(ex.map(one, frames * 2))
            el = time.perf_counter() - t0
        out["q0_host_fps"] = round(2 * n / el, 1)
        out["host_note"] = "host planes in and out per frame, 8 threads (PCIe-inclusive); never the value"
        clip.close()
    if not args.no_cpu_baseline:
        import oracle  # the CPU baseline only
        oracle.downsample_frame(oracle.downsample_params(quality=0), *frames[0], bits=10)  # warm-up
        t0 = time.perf_counter()
        for f in frames[1:5]:
            oracle.downsample_frame(oracle.downsample_params(quality=0), *f, bits=10)
        out["cpu_oracle_q0_fps"] = round(4 / (time.perf_counter() - t0), 2)
        out["cpu_oracle_note"] = "oracle C restatement, one host thread"
    del din, dout
    dev.close()
    torch.cuda.empty_cache()
    return out


def trecon(args):
    """TRecon (trad 1: two neighbours, block ME, temporal taps) on
    device-resident 4:2:0 P10 clips at 1080p and 4K, one frame at a time
    (each call waits for its frame): milliseconds per frame."""
    from lgcr_amd import abi, serve
    dev = abi.Device(0)
    out = {"workload": "TRecon YUV420P10 -> YUV444P10, lanczos3, trad 1, one frame at a time"}
    for name, (w, h) in (("1080p", (1920, 1080)), ("2160p", (3840, 2160))):
        n = 24
        clip = serve.ResidentClip(dev, serve.ClipSpec(name, "trecon", w, h, "420p10", n, 7), max_inflight=1,
                                  threads=min(16, os.cpu_count() or 1))
        clip.admit()
        for k in range(n):
            clip.run_blocking(k)
        t0 = time.perf_counter()
        for k in range(n):
            clip.run_blocking(k)
        out[f"{name}_ms_per_frame"] = round((time.perf_counter() - t0) / n * 1e3, 3)
        clip.release()
    dev.close()
    return out


# Non-headline workloads of the benchmark definition, each timed on
# device-resident frames with eight frames in flight, plus an isolated stage
# pass: (name, description, width, height, ss_w, ss_h, bits, out_w, out_h,
# Recon arguments, synthetic-noise arguments, frames)
WORKLOADS = (
    ("1080p420p10_algo4", "Recon 1920x1080 YUV420P10 -> YUV444P10, algo 4, lanczos3, bp 0.5",
     1920, 1080, 1, 1, 10, 1920, 1080, {"algo": 4, "bp": 0.5}, {}, 48),
    ("1080p420p10_strength0", "Recon 1920x1080 YUV420P10 -> YUV444P10, strength 0 (plain A/B control), bp 0.5",
     1920, 1080, 1, 1, 10, 1920, 1080, {"strength": 0.0, "bp": 0.5}, {}, 48),
    ("1080p420ps_to_2160p_jinc3_algo6", "Recon 1920x1080 YUV420PS -> 3840x2160 YUV444PS, jinc 3, algo 6",
     1920, 1080, 1, 1, 32, 3840, 2160, {"algo": 6, "kernel": "jinc", "taps": 3}, {}, 24),
    ("1080p420ps_to_2160p_lanczos4_algo6", "Recon 1920x1080 YUV420PS -> 3840x2160 YUV444PS, lanczos 4, algo 6",
     1920, 1080, 1, 1, 32, 3840, 2160, {"algo": 6, "kernel": "lanczos", "taps": 4}, {}, 24),
    ("1080p422p8_bm", "Recon 1920x1080 YUV422P8 (noise 3 LSB on all planes) -> YUV444P8, algo 2, bm, bp 0.5",
     1920, 1080, 1, 0, 8, 1920, 1080, {"bm": 1, "bp": 0.5}, {"luma_noise": 3.0, "chroma_noise": 3.0}, 48),
)


def workloads(args, only=None):
    """fps, per-stage device time and the dominant stage's HBM fraction for
    each entry of WORKLOADS (only: that entry alone)."""
    import numpy as np
    import torch
    from lgcr_amd import abi, synth
    dev = abi.Device(0)
    out = {}
    for name, desc, W, H, ssw, ssh, bits, ow, oh, kw, noise, n in WORKLOADS:
        if only is not None and name != only:
            continue
        bps = 4 if bits == 32 else (1 if bits <= 8 else 2)
        np_dt = {1: np.uint8, 2: np.uint16, 4: np.float32}[bps]
        tdt = {1: torch.uint8, 2: torch.int16, 4: torch.float32}[bps]
        frames = synth.clip(n, W, H, ss_w=ssw, ss_h=ssh, bits=bits, seed=3, scene_per_frame=True,
                            threads=min(16, os.cpu_count() or 1), **noise)
        din = [[torch.from_numpy(p.view(np.int16) if bps == 2 else p).to("cuda") for p in f] for f in frames]
        same = (ow, oh) == (W, H)
        dout = [[torch.empty((oh, ow), dtype=tdt, device="cuda") for _ in range(2 if same else 3)]
                for _ in frames]
        fin = [[(t.data_ptr(), t.stride(0) * bps) for t in f] for f in din]
        fout = [([(0, 0)] if same else []) + [(t.data_ptr(), t.stride(0) * bps) for t in o] for o in dout]
        fmt = abi.make_format(np_dt, W, H, ssw, ssh, bits=None if bits in (8, 16, 32) else bits)
        params = abi.default_params(width=ow, height=oh, **kw)
        clip = abi.Clip(dev, params, fmt, max_inflight=8)
        clip.recon_batch(fin, fout)
        torch.cuda.synchronize()
        steps = 2
        t0 = time.perf_counter()
        for _ in range(steps):
            clip.recon_batch(fin, fout)
        torch.cuda.synchronize()
        fps = steps * n / (time.perf_counter() - t0)
        clip.close()
        sclip = abi.Clip(dev, params, fmt, max_inflight=1)
        sclip.recon_batch(fin[:4], fout[:4])
        sclip.set_timing(True)
        sclip.recon_batch(fin, fout)
        torch.cuda.synchronize()
        stages = {k: v[0] / max(1, v[1]) for k, v in sclip.stage_times().items()}
        sclip.close()
        cw, ch = W >> ssw, H >> ssh
        nbytes = bps * (W * H + 2 * cw * ch + 2 * ow * oh + (0 if same else ow * oh))
        dom = max(stages, key=stages.get)
        out[name] = {
            "workload": desc, "frames": n, "inflight": 8, "fps": round(fps, 1),
            "bytes_per_frame": nbytes, "end_to_end_hbm_gbs": round(nbytes * fps / 1e9, 2),
            "stage_avg_us": {k: round(v, 2) for k, v in stages.items()},
            "device_us_per_frame": round(sum(stages.values()), 2),
            "dominant": {"stage": dom, "avg_us": round(stages[dom], 2),
                         "hbm_frac": round(nbytes / (stages[dom] * 1e-6) / 1e9 / HBM_PEAK_GBS, 5)},
        }
        del din, dout, fin, fout
        torch.cuda.empty_cache()
    dev.close()
    return out


def batched_launch(args):
    """Multi-frame launches (lgcr_clip_set_batching: every stage as one
    launch per group of frames) against one launch per frame, on
    device-resident frames: the 640x360 YUV420P8 workload with Recon
    defaults and the 1080p P10 benchmark workload.  Per frame: eight and
    sixteen workspaces, and sixteen workspaces replaying one captured graph
    per frame; batched: sixteen workspaces, groups of four frames on four
    streams, the tile pass as one wave per tile over a (tiles, frames) grid
    and as a persistent grid.  Each mode runs one warm-up and three timed
    passes over the frames."""
    import numpy as np
    import torch
    from lgcr_amd import abi, synth
    dev = abi.Device(0)
    out = {}
    cases = (("640x360_420p8", "Recon 640x360 YUV420P8 -> YUV444P8, defaults (algo 2, lanczos3, strength 0.8)",
              640, 360, 8, {}, 480),
             ("1080p420p10", "Recon 1920x1080 YUV420P10 -> YUV444P10, algo 2, lanczos3, strength 0.8, bp 0.5",
              1920, 1080, 10, {"bp": 0.5}, 96))
    for name, desc, W, H, bits, kw, n in cases:
        if args.batched_case and name != args.batched_case:
            continue
        bps = 1 if bits <= 8 else 2
        np_dt = np.uint8 if bps == 1 else np.uint16
        tdt = torch.uint8 if bps == 1 else torch.int16
        frames = synth.clip(n, W, H, bits=bits, seed=0, scene_per_frame=True, threads=min(16, os.cpu_count() or 1))
        din = [[torch.from_numpy(p.view(np.int16) if bps == 2 else p).to("cuda") for p in f] for f in frames]
        dout = [[torch.empty((H, W), dtype=tdt, device="cuda") for _ in range(2)] for _ in frames]
        fin = [[(t.data_ptr(), t.stride(0) * bps) for t in f] for f in din]
        fout = [[(0, 0)] + [(t.data_ptr(), t.stride(0) * bps) for t in o] for o in dout]
        fmt = abi.make_format(np_dt, W, H, bits=None if bits == 8 else bits)
        params = abi.default_params(**kw)
        res = {"workload": desc, "frames": n}
        for tag, inflight, batching, persistent, graphs in (("per_frame_8", 8, 0, False, False),
                                                           ("per_frame_16", 16, 0, False, False),
                                                           ("graphs_16", 16, 0, False, True),
                                                           ("batched_4x4_grid", 16, 4, False, False),
                                                           ("batched_4x4_persistent", 16, 4, True, False)):
            if args.batched_mode and tag != args.batched_mode:
                continue
            os.environ["LGCR_BATCH_PERSISTENT"] = "1" if persistent else "0"
            clip = abi.Clip(dev, params, fmt, max_inflight=inflight)
            os.environ.pop("LGCR_BATCH_PERSISTENT")
            clip.set_batching(batching)
            clip.set_graphs(graphs)
            clip.recon_batch(fin, fout)
            torch.cuda.synchronize()
            steps = 3
            t0 = time.perf_counter()
            for _ in range(steps):
                clip.recon_batch(fin, fout)
            torch.cuda.synchronize()
            res[f"{tag}_fps"] = round(steps * n / (time.perf_counter() - t0), 1)
            clip.close()
        out[name] = res
        del din, dout, fin, fout
        torch.cuda.empty_cache()
    dev.close()
    return out


def serving(args):
    """Clips of mixed resolution / format / filter with Poisson frame
    arrivals: continuous batching vs one blocking call per frame, at the
    offered rate and saturated (all requests at t=0)."""
    from lgcr_amd import abi, serve
    dev = abi.Device(0)
    specs = serve.make_mix(args.serve_clips, args.serve_frames, seed=0)
    clips = [serve.ResidentClip(dev, s, max_inflight=4, threads=min(16, os.cpu_count() or 1),
                                graphs=args.serve_graphs, keep_host=False) for s in specs]
    resident = sum(c.src_bytes + c.out_bytes for c in clips)
    rates = [float(r) for r in str(args.serve_rate).split(",") if r]
    out = {"clips": len(specs), "frames_per_clip": args.serve_frames, "offered_fps": rates,
           "poll_us": args.serve_poll_us, "max_inflight": args.serve_inflight, "max_work": args.serve_max_work,
           "graphs": args.serve_graphs,
           "mix": {k: sum(1 for s in specs if s.kind == k) for k in ("recon", "trecon")},
           "resident_frames_gb": round(resident / 2**30, 3)}
    loads = [(f"offered{r:g}", serve.poisson_requests(specs, r, seed=0)) for r in rates]
    loads.append(("saturated", serve.poisson_requests(specs, float("inf"), seed=0)))
    # untimed pass: admits every clip (device state, geometry, workspaces) so
    # that neither mode's latencies include first-request set-up
    serve.serve(clips, loads[-1][1], mode="continuous", max_inflight=args.serve_inflight)
    for name, reqs in loads:
        for mode in ("continuous", "per_frame"):
            st = serve.serve(clips, reqs, mode=mode, poll_sleep=args.serve_poll_us * 1e-6,
                             max_inflight=args.serve_inflight, max_work=args.serve_max_work)
            out[f"{name}_{mode}"] = {"fps": round(st["fps"], 1), "p50_ms": round(st["p50_ms"], 3),
                                     "p99_ms": round(st["p99_ms"], 3)}
            out["peak_hbm_gb"] = round(max(out.get("peak_hbm_gb", 0.0),
                                           st["peak_reserved_gb"] + resident / 2**30), 3)
    for c in clips:
        c.release()
    dev.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=240)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--bits", type=int, default=10)
    ap.add_argument("--algo", type=int, default=2)
    ap.add_argument("--bp", type=float, default=0.5)
    ap.add_argument("--strength", type=float, default=0.8)
    ap.add_argument("--inflight", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-clips", type=int, default=64)
    ap.add_argument("--serve-frames", type=int, default=120)
    ap.add_argument("--serve-rate", default="600,1500", help="offered frame rates (comma-separated)")
    ap.add_argument("--no-workloads", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-downsample", action="store_true")
    ap.add_argument("--no-trecon", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-case", default=None, help="multi-frame section: one workload (640x360_420p8, 1080p420p10)")
    ap.add_argument("--batched-mode", default=None,
                    help="multi-frame section: one launch mode (per_frame_8, per_frame_16, graphs_16, "
                         "batched_4x4_grid, batched_4x4_persistent)")
    ap.add_argument("--serve-poll-us", type=float, default=20.0)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--serve-graphs", action="store_true")
    ap.add_argument("--serve-max-work", type=float, default=None,
                    help="continuous mode: relative work allowed in flight (1080p Recon = 1) unless idle")
    ap.add_argument("--no-host-path", action="store_true")
    ap.add_argument("--graphs", action="store_true")
    ap.add_argument("--batching", type=int, default=0,
                    help="frames per multi-frame fused tile launch (0: one launch per frame)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--section", default=None,
                    help="run one section alone and print its JSON (for profiling): downsample, trecon, serving, "
                         "or a workload name")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(Recon is data-parallel over frames: see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    if args.section:
        names = [w[0] for w in WORKLOADS]
        if args.section in ("downsample", "trecon", "serving", "batched"):
            out = {"downsample": downsample, "trecon": trecon, "serving": serving,
                   "batched": batched_launch}[args.section](args)
        elif args.section in names:
            out = workloads(args, only=args.section)
        else:
            print(f"bench.py: unknown section {args.section!r} (downsample, trecon, serving, batched, "
                  f"{', '.join(names)})",
                  file=sys.stderr)
            sys.exit(2)
        print(json.dumps({"section": args.section, **out}))
        return

    import numpy as np
    import torch
    from lgcr_amd import abi, synth

    W, H, bits = args.width, args.height, args.bits
    cw, ch = W // 2, H // 2
    bps = 4 if bits == 32 else (1 if bits <= 8 else 2)   # bits 32: YUV420PS
    frames = synth.clip(args.frames, W, H, bits=bits, seed=0, scene_per_frame=True,
                        threads=min(16, os.cpu_count() or 1))
    tdt = {1: torch.uint8, 2: torch.int16, 4: torch.float32}[bps]
    np_dt = {1: np.uint8, 2: np.uint16, 4: np.float32}[bps]

    def to_dev(a):
        return torch.from_numpy(a.view(np.int16) if bps == 2 else a).to("cuda")

    dev_in = [[to_dev(p) for p in f] for f in frames]
    dev_out = [[torch.empty((H, W), dtype=tdt, device="cuda") for _ in range(2)] for _ in frames]
    fin = [[(t.data_ptr(), t.stride(0) * bps) for t in f] for f in dev_in]
    fout = [[(0, 0)] + [(t.data_ptr(), t.stride(0) * bps) for t in o] for o in dev_out]
    torch.cuda.synchronize()

    dev = abi.Device(0)
    params = abi.default_params(algo=args.algo, bp=args.bp, strength=args.strength)
    clip = abi.Clip(dev, params, abi.make_format(np_dt, W, H, bits=None if bits == 32 else bits),
                    max_inflight=args.inflight)
    clip.set_graphs(args.graphs)
    clip.set_batching(args.batching)
    # tile shapes of the tunable kernels for this frame size (autotuner)
    tuning = clip.autotune_device(fin[0]) if not args.no_autotune else clip.tuning()
    for _ in range(args.warmup):
        clip.recon_batch(fin, fout)
    torch.cuda.synchronize()

    # timed region: the production configuration (no stage events; frame
    # graphs only with --graphs)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        clip.recon_batch(fin, fout)
    torch.cuda.synchronize()
    t1 = time.perf_counter()

    # per-stage device time: the same steps again on a one-workspace clip
    # (same tile shapes) with HIP events recorded around each stage on the
    # stream that runs it, so every kernel's duration is its own, not shared
    # with other frames' kernels (events need the direct launches)
    sclip = abi.Clip(dev, params, abi.make_format(np_dt, W, H, bits=None if bits == 32 else bits), max_inflight=1)
    sclip.set_tuning(*tuning)
    sclip.recon_batch(fin[:8], fout[:8])
    sclip.set_timing(True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        sclip.recon_batch(fin, fout)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    stages = sclip.stage_times()
    sclip.close()

    elapsed = t1 - t0
    n = args.frames * args.steps
    fps = n / elapsed
    per_frame_bytes = algorithmic_bytes(W, H, cw, ch, bps)
    dominant = max(stages.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_us, dom_n) = dominant
    dom_avg_us = dom_us / max(1, dom_n)
    achieved = per_frame_bytes / (dom_avg_us * 1e-6) / 1e9
    stage_avg = {k: round(v[0] / max(1, v[1]), 2) for k, v in stages.items()}

    traffic = valu = None
    if (W, H, bits, args.algo) == (1920, 1080, 10, 2) and dom_name == "ApplyGuidedCorrection":
        traffic = measured_traffic("recon_1080p420p10_algo2")
        valu = measured_valu("recon_1080p420p10_algo2")
    # attainable time of the dominant kernel: the larger of its HBM time and
    # its VALU time at the non-packed lane-op peak
    t_min = max(per_frame_bytes / (HBM_PEAK_GBS * 1e9), (valu or 0) / VALU_PEAK_LANE_OPS)
    algo_ops = algorithmic_valu(clip, frames) if dom_name == "ApplyGuidedCorrection" else None
    fmt_name = "PS" if bits == 32 else f"P{bits}"
    result = {
        "metric": METRIC,
        "value": round(fps, 2),
        "unit": "fps",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1000.0, 3),
        "frame_graphs": args.graphs,
        "batching": args.batching,
        "tuning": {"tensor_tile_rows": tuning[0], "gate_tile_cols": tuning[1], "autotuned": not args.no_autotune},
        "stage_pass": {"inflight": 1, "fps": round(n / (t3 - t2), 2)},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulators)",
        "data": "synthetic (polygon scenes, seed 0, one scene per frame)",
        "config": {"workload": f"Recon {W}x{H} YUV420{fmt_name} -> YUV444{fmt_name}, algo {args.algo}, "
                               f"lanczos3, strength {args.strength}, sparse 1, bp {args.bp}",
                   "frames_per_step": args.frames, "inflight_workspaces": args.inflight},
        "end_to_end_hbm_gbs": round(per_frame_bytes * fps / 1e9, 2),
        "end_to_end_hbm_frac": round(per_frame_bytes * fps / 1e9 / HBM_PEAK_GBS, 5),
        "stage_avg_us": stage_avg,
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "bytes_per_launch": per_frame_bytes,
            "avg_launch_us": round(dom_avg_us, 2),
            "valu_lane_ops_per_launch": valu,
            "valu_peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
            "attainable_frac": round(t_min / (dom_avg_us * 1e-6), 4) if valu else None,
            "algorithmic_lane_ops_per_launch": round(algo_ops) if algo_ops else None,
            "algorithmic_valu_frac": (round(algo_ops / VALU_PEAK_LANE_OPS / (dom_avg_us * 1e-6), 4)
                                      if algo_ops else None),
            "lane_op_economy": round(algo_ops / valu, 4) if (algo_ops and valu) else None,
        },
    }
    clip.close()
    if not args.no_host_path and bits != 32:
        hclip = abi.Clip(dev, params, abi.make_format(np_dt, W, H, bits=bits), max_inflight=args.inflight)
        host_path(hclip, frames, 8, args.inflight)  # warm-up
        n_host = min(len(frames), 96)
        result["host_buffers"] = {
            "fps": round(host_path(hclip, frames, n_host, args.inflight), 1), "frames": n_host,
            "threads": args.inflight,
            "note": "host planes in and out per frame (PCIe-inclusive; the VapourSynth plugin boundary)"}
        hclip.close()
    dev.close()
    del dev_in, dev_out, fin, fout
    torch.cuda.empty_cache()

    if not args.no_workloads:
        result["workloads"] = workloads(args)

    if not args.no_batched:
        result["batched_launch"] = batched_launch(args)

    if not args.no_serve:
        result["serving"] = serving(args)

    if not args.no_downsample:
        result["downsample"] = downsample(args)

    if not args.no_trecon:
        result["trecon"] = trecon(args)

    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = frames[: max(1, min(len(frames), 2 * threads))]
        kw = {"algo": args.algo, "bp": args.bp, "strength": args.strength}
        v, nf, el = cpu_baseline(sample, kw, bits, args.cpu_seconds, threads)
        result["cpu_baseline"] = {
            "value": round(v, 3), "unit": "fps", "cores": threads, "kind": "port",
            "sample": f"{nf} frames of the same workload in {el:.1f} s on {threads} threads "
                      f"(oracle C restatement, scalar, one frame per thread)",
        }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
