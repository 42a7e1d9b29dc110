"""Serving loop host logic on the CPU: request generation, neighbour choice,
and the continuous-batching scheduler driven by a simulated backend
(admission by budget, busy workspaces, completion polling)."""
import time

import pytest

from lgcr_amd import abi, serve


def test_mix_and_poisson_requests():
    specs = serve.make_mix(12, 7, seed=3)
    assert len(specs) == 12
    assert {s.kind for s in specs} <= {"recon", "trecon"}
    assert all((s.width, s.height) in serve.RESOLUTIONS and s.fmt in serve.FORMATS for s in specs)
    reqs = serve.poisson_requests(specs, 600.0, seed=3)
    assert len(reqs) == 12 * 7
    assert all(a.t <= b.t for a, b in zip(reqs, reqs[1:]))
    for i in range(12):
        assert [r.frame for r in reqs if r.clip == i] == list(range(7))
    # mean inter-arrival close to 1/600 s
    assert abs(reqs[-1].t / len(reqs) - 1 / 600) < 0.5 / 600
    sat = serve.poisson_requests(specs, float("inf"))
    assert all(r.t == 0.0 for r in sat)


def test_temporal_frame_numbers_match_reference_rule():
    assert serve.temporal_frame_numbers(0, 9, 1) == [1]
    assert serve.temporal_frame_numbers(5, 9, 1) == [4, 6]
    assert serve.temporal_frame_numbers(9, 9, 2) == [7, 8]
    assert serve.temporal_frame_numbers(0, 0, 3) == []


class FakeDevice:
    def __init__(self, budget):
        self.budget = budget
        self.used = 0
        self.peak = 0

    def reserved(self):
        return self.used


class FakeClip:
    """Backend stand-in: `slots` workspaces, each frame takes `cost` seconds."""

    def __init__(self, dev, size, slots=2, cost=0.002):
        self.device = dev
        self.size = size
        self.slots = slots
        self.cost = cost
        self.free_out = list(range(slots))
        self.clip = None
        self.rejected_at = -1
        self.running = {}
        self.order = []
        self.concurrent_peak = 0

    def admit(self):
        if self.clip is not None:
            return True
        if self.device.used + self.size > self.device.budget:
            return False
        self.device.used += self.size
        self.clip = object()
        return True

    def release(self):
        if self.clip is not None:
            self.device.used -= self.size
            self.clip = None

    def submit(self, n):
        if not self.free_out:
            return None
        slot = self.free_out.pop()
        t = object()
        self.running[t] = time.perf_counter() + self.cost
        self.order.append(n)
        self.concurrent_peak = max(self.concurrent_peak, len(self.running))
        return t, slot

    def poll(self, t):
        return time.perf_counter() >= self.running[t]

    def finish(self, t, slot):
        del self.running[t]
        self.free_out.append(slot)

    def run_blocking(self, n, slot=0):
        self.order.append(n)
        time.sleep(self.cost)

    def output_host(self, slot, n):
        return ("frame", n)


def test_continuous_batching_overlaps_and_keeps_clip_order():
    dev = FakeDevice(budget=10)
    clips = [FakeClip(dev, 3, slots=2) for _ in range(4)]
    specs = [serve.ClipSpec(f"c{i}", "recon", 64, 64, "420p8", 6, i) for i in range(4)]
    reqs = serve.poisson_requests(specs, float("inf"))
    keep = {}
    st = serve.serve(clips, reqs, mode="continuous", keep=keep)
    assert st["frames"] == 24 and len(keep) == 24
    for c in clips:
        assert c.order == list(range(6))       # a clip's frames run in order
    # budget 10 admits three clips of 3 at once; the fourth waits until a
    # finished clip releases its state; frames of a clip overlap
    assert max(c.concurrent_peak for c in clips) == 2
    assert dev.used <= 10
    # overlap: 24 frames of 2 ms on up to 6 concurrent slots finish well
    # under the 48 ms a serial loop needs
    assert st["seconds"] < 0.040


def test_admission_waits_for_release():
    dev = FakeDevice(budget=4)
    clips = [FakeClip(dev, 3, slots=1) for _ in range(3)]   # only one fits at a time
    specs = [serve.ClipSpec(f"c{i}", "recon", 64, 64, "420p8", 3, i) for i in range(3)]
    reqs = serve.poisson_requests(specs, float("inf"))
    st = serve.serve(clips, reqs, mode="continuous")
    assert st["frames"] == 9
    assert st["peak_reserved_gb"] * 2**30 <= 4


def test_per_frame_mode_is_serial():
    dev = FakeDevice(budget=100)
    clips = [FakeClip(dev, 1, slots=1, cost=0.001) for _ in range(2)]
    specs = [serve.ClipSpec(f"c{i}", "recon", 64, 64, "420p8", 4, i) for i in range(2)]
    reqs = serve.poisson_requests(specs, float("inf"))
    st = serve.serve(clips, reqs, mode="per_frame")
    assert st["frames"] == 8 and st["seconds"] >= 0.008
    assert st["p99_ms"] >= st["p50_ms"] > 0


@pytest.mark.parametrize("mode", ["continuous", "per_frame"])
def test_clip_larger_than_budget_raises(mode):
    # a budget smaller than one clip's workspace: nothing can ever make room,
    # so both modes refuse instead of waiting forever
    dev = FakeDevice(budget=2)
    clips = [FakeClip(dev, 1, slots=1, cost=0.001), FakeClip(dev, 3, slots=1, cost=0.001)]
    specs = [serve.ClipSpec(f"c{i}", "recon", 64, 64, "420p8", 2, i) for i in range(2)]
    reqs = serve.poisson_requests(specs, float("inf"))
    with pytest.raises(RuntimeError, match="does not fit the device budget alone"):
        serve.serve(clips, reqs, mode=mode)


class WaitingClip(FakeClip):
    """FakeClip with a native-style wait: sleeps until the first of the
    tickets completes or the timeout passes."""
    waits = []

    @staticmethod
    def wait_any(tickets, timeout_s):
        WaitingClip.waits.append(timeout_s)
        owner = {t: c for c in WaitingClip.all for t in c.running}
        ends = [owner[t].running[t] for t in tickets]
        time.sleep(max(0.0, min(min(ends) - time.perf_counter(), timeout_s)))


def test_continuous_uses_clip_wait_any():
    """The continuous loop blocks in the clips' wait_any (bounded by the next
    arrival) instead of napping, and still serves every frame in order."""
    dev = FakeDevice(budget=10)
    WaitingClip.waits = []
    clips = [WaitingClip(dev, 3, slots=2) for _ in range(3)]
    WaitingClip.all = clips
    reqs = [serve.Request(0.001 * k, k % 3, k // 3) for k in range(12)]
    keep = {}
    st = serve.serve(clips, reqs, mode="continuous", keep=keep)
    assert st["frames"] == 12 and len(keep) == 12
    assert WaitingClip.waits and all(0.0 <= w <= 5e-3 for w in WaitingClip.waits)
    for c in clips:
        assert c.order == sorted(c.order)


class FailingClip(FakeClip):
    """FakeClip whose backend call fails for the frames in `bad`."""

    def __init__(self, *a, bad=(), **k):
        super().__init__(*a, **k)
        self.bad = set(bad)

    def submit(self, n):
        if n in self.bad:
            raise abi.LgcrError(abi.LGCR_BACKEND_FAILURE, "kernel launch failed")
        return super().submit(n)

    def run_blocking(self, n, slot=0):
        if n in self.bad:
            raise abi.LgcrError(abi.LGCR_BACKEND_FAILURE, "kernel launch failed")
        super().run_blocking(n, slot)


@pytest.mark.parametrize("mode", ["continuous", "per_frame"])
def test_failed_frame_does_not_stop_serving(mode):
    """A frame whose backend call fails is reported and skipped; the clip's
    later frames and the other clips are still served."""
    dev = FakeDevice(budget=10)
    clips = [FailingClip(dev, 3, slots=2, bad={1}), FakeClip(dev, 3, slots=2)]
    reqs = [serve.Request(0.0005 * k, k % 2, k // 2) for k in range(8)]
    keep, failures = {}, []
    st = serve.serve(clips, reqs, mode=mode, keep=keep, failures=failures)
    assert st["frames"] == 7 and st["failed"] == 1
    assert [(r.clip, r.frame) for r, _ in failures] == [(0, 1)]
    assert "kernel launch failed" in failures[0][1]
    assert set(keep) == {(c, f) for c in (0, 1) for f in range(4)} - {(0, 1)}
