"""Serving loop: frame requests from many clips of different resolutions and
formats arrive over time and run on one device with continuous batching.

* Every clip's source frames are uploaded once and stay resident in HBM
  (TRecon reads its neighbours from there).
* A clip's device state (geometry tables + per-frame workspaces) is created
  when its first request is served, admitted against the device's HBM budget
  (``LGCR_REJECTED`` keeps the clip waiting until another clip finishes and
  releases its workspaces), and destroyed after its last frame.
* Continuous batching: one scheduler thread submits every queued request
  whose clip has a free workspace (``lgcr_*_submit`` returns ``LGCR_BUSY``
  otherwise and the scheduler moves on to other clips), polls the tickets of
  frames on the device, and admits new arrivals while others run.  Frames of
  different clips overlap on the device on separate HIP streams.
* The per-frame baseline serves the same arrivals one blocking call at a
  time (``lgcr_recon_frame`` / ``lgcr_trecon_frame``).

The reference's closest analogue is its VapourSynth benchmark harness
(test/benchmark.py:322-440): a static source, ``inflight`` outstanding
requests, completion-latency percentiles.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import abi, synth

FORMATS = {"420p8": (1, 1, 8), "420p10": (1, 1, 10), "422p10": (1, 0, 10)}
RESOLUTIONS = [(720, 480), (1280, 720), (1920, 1080), (3840, 2160)]


@dataclass
class ClipSpec:
    name: str
    kind: str            # "recon" or "trecon"
    width: int
    height: int
    fmt: str             # key of FORMATS
    frames: int
    seed: int


@dataclass(order=True)
class Request:
    t: float                     # arrival time, seconds from the start
    clip: int = field(compare=False)
    frame: int = field(compare=False)


def make_mix(n_clips, frames, seed=0, resolutions=RESOLUTIONS, formats=tuple(FORMATS), kinds=("recon", "trecon")):
    """Clip specs with resolution, format and filter drawn uniformly."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_clips):
        w, h = resolutions[rng.integers(len(resolutions))]
        out.append(ClipSpec(f"clip{i}", kinds[rng.integers(len(kinds))], int(w), int(h),
                            formats[rng.integers(len(formats))], frames, int(seed * 1000 + i)))
    return out


def poisson_requests(specs, rate, seed=0):
    """Poisson arrivals at `rate` frames/s in total; each arrival goes to a
    random clip with frames left, which plays its frames in order.  rate=inf
    puts every request at t=0 (saturation)."""
    rng = np.random.default_rng(seed + 1)
    left = [s.frames for s in specs]
    nxt = [0] * len(specs)
    t, out = 0.0, []
    total = sum(left)
    for _ in range(total):
        if math.isfinite(rate):
            t += float(rng.exponential(1.0 / rate))
        live = [i for i, n in enumerate(left) if n > 0]
        i = live[rng.integers(len(live))]
        out.append(Request(t, i, nxt[i]))
        nxt[i] += 1
        left[i] -= 1
    return out


def temporal_frame_numbers(n, n_last, trad):
    """TRecon neighbours: n-trad..n+trad clamped, centre and duplicates
    removed (plugin.cpp:729-738)."""
    out = []
    for k in range(-trad, trad + 1):
        fn = min(max(n + k, 0), n_last)
        if fn != n and fn not in out:
            out.append(fn)
    return out


class ResidentClip:
    """A clip's source frames in HBM, its output buffers and its backend
    state (created on admission)."""

    def __init__(self, device, spec: ClipSpec, max_inflight=4, threads=8, graphs=False, keep_host=True):
        import torch
        self.spec = spec
        self.graphs = graphs  # Recon frames replay captured frame graphs
        self.device = device
        ssw, ssh, bits = FORMATS[spec.fmt]
        self.ssw, self.ssh, self.bits = ssw, ssh, bits
        self.bps = 1 if bits <= 8 else 2
        self.cw, self.ch = spec.width >> ssw, spec.height >> ssh
        frames = synth.clip(spec.frames, spec.width, spec.height, ss_w=ssw, ss_h=ssh, bits=bits, seed=spec.seed,
                            motion=True, threads=threads)
        view = (lambda a: a.view(np.int16)) if self.bps == 2 else (lambda a: a)
        self.src = [[torch.from_numpy(view(p)).to("cuda") for p in f] for f in frames]
        # host copies serve output_host (the shared Recon luma) and checks;
        # keep_host=False drops them once the frames are resident
        self.host_src = frames if keep_host else None
        self.temporal = spec.kind == "trecon"
        self.max_inflight = max_inflight
        dt = torch.uint8 if self.bps == 1 else torch.int16
        w, h = spec.width, spec.height
        # one output frame per workspace slot: 4:4:4 at the input size
        self.out = [[torch.empty((h, w), dtype=dt, device="cuda") for _ in range(3)] for _ in range(max_inflight)]
        self.free_out = list(range(max_inflight))
        self.clip = None
        self.rejected_at = -1     # release generation of the last refused admission
        # relative device work of one frame (1080p Recon = 1): TRecon also
        # converts and block-matches its two neighbours
        self.cost = spec.width * spec.height / (1920 * 1080) * (4.0 if self.temporal else 1.0)
        self.src_bytes = sum(t.numel() * t.element_size() for f in self.src for t in f)
        self.out_bytes = sum(t.numel() * t.element_size() for f in self.out for t in f)

    def frame(self, planes, share_luma=False):
        """Device frame view; share_luma leaves plane 0 NULL (same-size Recon
        shares the source luma, plugin.cpp:265-269)."""
        bps = self.bps
        ptrs = [t.data_ptr() for t in planes]
        strides = [t.stride(0) * bps for t in planes]
        if share_luma:
            ptrs[0], strides[0] = None, 0
        return abi.Frame((C.c_void_p * 3)(*ptrs), (C.c_ssize_t * 3)(*strides), abi.MEMORY_DEVICE)

    def params(self):
        return abi.trecon_params() if self.temporal else abi.default_params()

    def admit(self):
        """Create the backend state; False when the device budget refuses it."""
        if self.clip is not None:
            return True
        fmt = abi.make_format(np.uint8 if self.bps == 1 else np.uint16, self.spec.width, self.spec.height,
                              self.ssw, self.ssh, self.bits if self.bits not in (8, 16) else None)
        for attempt in range(2):
            try:
                self.clip = abi.Clip(self.device, self.params(), fmt, max_inflight=self.max_inflight,
                                     temporal=self.temporal)
                break
            except abi.LgcrError as e:
                # a failed allocation is retried once; refused by the budget
                # or out of memory again, the clip waits for memory to be
                # released
                if e.status == abi.LGCR_OUT_OF_MEMORY and attempt == 0:
                    continue
                if e.status in (abi.LGCR_REJECTED, abi.LGCR_OUT_OF_MEMORY):
                    return False
                raise
        if self.graphs:
            self.clip.set_graphs(True)
        return True

    def release(self):
        if self.clip is not None:
            self.clip.close()
            self.clip = None

    def _nbrs(self, n):
        ids = temporal_frame_numbers(n, self.spec.frames - 1, 1)
        arr = (abi.Frame * max(1, len(ids)))()
        for k, i in enumerate(ids):
            arr[k] = self.frame(self.src[i])
        return arr, len(ids)

    def submit(self, n):
        """Queue frame n; returns (ticket, out_slot) or None when busy."""
        if not self.free_out:
            return None
        slot = self.free_out[-1]
        fin, fout = self.frame(self.src[n]), self.frame(self.out[slot], share_luma=not self.temporal)
        t = C.c_void_p()
        L = abi.lib()
        if self.temporal:
            nb, nn = self._nbrs(n)
            st = L.lgcr_trecon_submit(self.clip.handle, C.byref(fin), nb, nn, math.nan, math.nan, C.byref(fout),
                                      C.byref(t))
        else:
            st = L.lgcr_recon_submit(self.clip.handle, C.byref(fin), math.nan, math.nan, C.byref(fout), C.byref(t))
        if st == abi.LGCR_BUSY:
            return None
        abi.check(st)
        self.free_out.pop()
        return t, slot

    def poll(self, ticket):
        """True when a submitted frame is complete."""
        return abi.lib().lgcr_ticket_query(ticket) == abi.LGCR_OK

    @staticmethod
    def wait_any(tickets, timeout_s):
        """Wait up to timeout_s for any of the tickets to complete (spinning
        on their completion events in native code, GIL released)."""
        arr = (C.c_void_p * len(tickets))(*tickets)
        idx = C.c_int(-1)
        st = abi.lib().lgcr_ticket_wait_any(arr, len(tickets), max(timeout_s, 0.0) * 1e6, C.byref(idx))
        if st != abi.LGCR_BUSY:
            abi.check(st)

    def finish(self, ticket, slot):
        """Retire a completed ticket: workspace and output slot become free."""
        abi.check(abi.lib().lgcr_ticket_wait(ticket))
        self.free_out.append(slot)

    def run_blocking(self, n, slot=0):
        fin, fout = self.frame(self.src[n]), self.frame(self.out[slot], share_luma=not self.temporal)
        L = abi.lib()
        if self.temporal:
            nb, nn = self._nbrs(n)
            abi.check(L.lgcr_trecon_frame(self.clip.handle, C.byref(fin), nb, nn, math.nan, math.nan,
                                          C.byref(fout), None))
        else:
            abi.check(L.lgcr_recon_frame(self.clip.handle, C.byref(fin), math.nan, math.nan, C.byref(fout), None))

    def output_host(self, slot, n):
        """Copy the output of `slot` to host numpy planes (Y shared for Recon)."""
        outs = []
        for i, t in enumerate(self.out[slot]):
            if i == 0 and not self.temporal:
                outs.append(self.host_src[n][0].copy())
                continue
            a = t.cpu().numpy()
            outs.append(a.view(np.uint16) if self.bps == 2 else a)
        return outs


def _percentiles(lat):
    if not lat:
        return 0.0, 0.0
    a = np.sort(np.asarray(lat))
    return float(np.percentile(a, 50) * 1e3), float(np.percentile(a, 99) * 1e3)


def _wait_inflight(clips, inflight, timeout_s, poll_sleep):
    """Block until a frame in flight completes or timeout_s passes: the
    clips' native wait when they have one, else a poll_sleep nap."""
    waiter = getattr(clips[inflight[0][1].clip], "wait_any", None)
    if waiter is None:
        time.sleep(min(poll_sleep, max(timeout_s, 0.0)))
    else:
        waiter([tk for tk, _, _ in inflight], timeout_s)


def serve(clips, requests, mode="continuous", keep=None, poll_sleep=20e-6, max_inflight=4, lat_out=None,
          max_work=None, failures=None):
    """Serve `requests` (sorted by arrival) on `clips` (ResidentClip list).

    mode "continuous": non-blocking submissions across clips (at most
    `max_inflight` frames on the device at once and, with `max_work`, at
    most that much relative work (`ResidentClip.cost`) unless the device is
    idle; oldest request first),
    completions polled; mode "per_frame": one blocking call at a time.
    A clip whose frames are all served keeps its device state until another
    clip is refused admission (releasing frees memory but synchronises the
    device).  keep: optional dict filled with {(clip, frame): host planes}
    for correctness checks.  lat_out: optional list that receives
    (request, latency in s) per served frame.  A frame whose backend call
    fails is not served: it goes to `failures` (a list of (request, message),
    when given) and the other frames and clips go on.  Returns a stats dict (fps, p50/p99 latency in
    ms, peak reserved HBM)."""
    remaining = [0] * len(clips)
    for r in requests:
        remaining[r.clip] += 1
    dev = clips[0].device if clips else None
    lat, peak = [], 0
    failed = [] if failures is None else failures
    t0 = time.perf_counter()
    i = 0
    if mode == "per_frame":
        for r in requests:
            now = time.perf_counter() - t0
            if r.t > now:
                time.sleep(r.t - now)
            c = clips[r.clip]
            if not c.admit():
                # make room: drop the device state of clips that are done
                for k, o in enumerate(clips):
                    if remaining[k] == 0:
                        o.release()
                if not c.admit():
                    raise RuntimeError("per-frame serving: clip does not fit the device budget alone")
            try:
                c.run_blocking(r.frame)
            except abi.LgcrError as e:
                failed.append((r, str(e)))
                remaining[r.clip] -= 1
                continue
            lat.append(time.perf_counter() - t0 - r.t)
            if lat_out is not None:
                lat_out.append((r, lat[-1]))
            peak = max(peak, dev.reserved())
            if keep is not None:
                keep[(r.clip, r.frame)] = c.output_host(0, r.frame)
            remaining[r.clip] -= 1
        el = time.perf_counter() - t0
        p50, p99 = _percentiles(lat)
        return {"frames": len(lat), "failed": len(failed), "seconds": el, "fps": len(lat) / el, "p50_ms": p50,
                "p99_ms": p99, "peak_reserved_gb": peak / 2**30}

    pending = deque()
    inflight = []   # (ticket, request, slot)
    work = 0.0      # relative work of the frames in flight
    gen = 0         # bumped whenever a clip releases its device state
    while i < len(requests) or pending or inflight:
        now = time.perf_counter() - t0
        while i < len(requests) and requests[i].t <= now:
            pending.append(requests[i])
            i += 1
        progressed = False
        # completions first: they free workspaces and output slots
        still = []
        for tk, r, slot in inflight:
            c = clips[r.clip]
            if c.poll(tk):
                c.finish(tk, slot)
                work -= c.cost
                lat.append(time.perf_counter() - t0 - r.t)
                if lat_out is not None:
                    lat_out.append((r, lat[-1]))
                if keep is not None:
                    keep[(r.clip, r.frame)] = c.output_host(slot, r.frame)
                remaining[r.clip] -= 1
                if remaining[r.clip] == 0 and any(o.clip is None and o.rejected_at >= 0 for o in clips):
                    # a clip is waiting for admission: free this one's memory
                    c.release()
                    gen += 1
                progressed = True
            else:
                still.append((tk, r, slot))
        inflight = still
        # submit queued requests in arrival order; a busy clip does not block
        # the others (continuous batching), but a clip's frames stay in order
        blocked = set()
        over_budget = False
        rest = deque()
        while pending:
            r = pending.popleft()
            c = clips[r.clip]
            if len(inflight) >= max_inflight or over_budget:
                rest.append(r)
                continue
            if max_work is not None and inflight and work + c.cost > max_work:
                # the oldest request that does not fit holds the later ones
                # back, so small frames cannot starve a large one
                over_budget = True
                rest.append(r)
                continue
            if r.clip in blocked or (c.clip is None and c.rejected_at == gen):
                blocked.add(r.clip)
                rest.append(r)
                continue
            if not c.admit():
                # admission refused by the HBM budget: release clips whose
                # frames are all served, else retry after the next release
                freed = False
                for k, o in enumerate(clips):
                    if remaining[k] == 0 and o.clip is not None:
                        o.release()
                        freed = True
                if freed:
                    gen += 1
                if not (freed and c.admit()):
                    if all(o.clip is None for k, o in enumerate(clips) if k != r.clip):
                        # no other clip holds device state, so no release can
                        # ever make room for this one
                        raise RuntimeError("continuous serving: clip does not fit the device budget alone")
                    c.rejected_at = gen
                    blocked.add(r.clip)
                    rest.append(r)
                    continue
            try:
                res = c.submit(r.frame)
            except abi.LgcrError as e:
                # the frame fails; the clip and the other clips go on
                failed.append((r, str(e)))
                remaining[r.clip] -= 1
                progressed = True
                continue
            if res is None:
                blocked.add(r.clip)
                rest.append(r)
                continue
            inflight.append((res[0], r, res[1]))
            work += c.cost
            progressed = True
        pending = rest
        peak = max(peak, dev.reserved())
        if not progressed:
            if inflight:
                # until a frame completes or the next request arrives
                nxt = requests[i].t - (time.perf_counter() - t0) if i < len(requests) else 5e-3
                _wait_inflight(clips, inflight, max(0.0, min(nxt, 5e-3)), poll_sleep)
            elif pending and i >= len(requests):
                time.sleep(poll_sleep)
            elif i < len(requests):
                wait = requests[i].t - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 1e-3))
    el = time.perf_counter() - t0
    p50, p99 = _percentiles(lat)
    return {"frames": len(lat), "failed": len(failed), "seconds": el, "fps": len(lat) / el, "p50_ms": p50,
            "p99_ms": p99, "peak_reserved_gb": peak / 2**30}
