"""Failure injection through LGCR_INJECT_FAULT ("launch:<n>" fails the n-th
frame launch, "alloc:<n>" the n-th clip workspace allocation, each once):
the failing frame raises the filter's frame error while the same clip's
next frame, the other clips and the serving loop go on.  The reference's
executor reports a failed stage per frame (StageStatus,
src/cuda_backend.cpp:198-240); this backend has no CPU fallback, so the
frame errors instead."""
import numpy as np
import pytest

import oracle
from lgcr_amd import abi, serve, synth
from lgcr_amd.plugin import INTEGER, YUV, Error, clip_from_planes, core, query_video_format

pytestmark = pytest.mark.gpu

P420 = query_video_format(YUV, INTEGER, 10, 1, 1)
P444 = query_video_format(YUV, INTEGER, 10, 0, 0)


def source(w, h, n, seed, ss=(1, 1)):
    frames = [synth.clip_frame(w, h, ss_w=ss[0], ss_h=ss[1], bits=10, seed=seed, frame=k, motion=True)
              for k in range(n)]
    return frames, clip_from_planes(frames, P420 if ss == (1, 1) else P444)


def recon_ref(w, h, planes):
    return oracle.recon_frame(oracle.make_params(w, h), *planes, bits=10)


def same(frame, ref):
    return all(np.array_equal(np.asarray(frame[i]), ref[i]) for i in range(3))


def test_launch_fault_fails_one_frame(device, monkeypatch):
    fa, a = source(96, 64, 3, seed=3)
    fb, b = source(128, 72, 2, seed=4)
    ra, rb = core.lgcr.Recon(a), core.lgcr.Recon(b)
    assert same(ra.get_frame(0), recon_ref(96, 64, fa[0]))
    assert same(rb.get_frame(0), recon_ref(128, 72, fb[0]))
    monkeypatch.setenv("LGCR_INJECT_FAULT", "launch:1")
    with pytest.raises(Error, match=r"^LGCR: .*kernel launch failed: injected"):
        ra.get_frame(1)
    # fired once: the other clip and the same clip's next frames are served
    assert same(rb.get_frame(1), recon_ref(128, 72, fb[1]))
    assert same(ra.get_frame(2), recon_ref(96, 64, fa[2]))
    assert same(ra.get_frame(1), recon_ref(96, 64, fa[1]))


def test_launch_fault_second_check(device, monkeypatch):
    """launch:2 lets the first launch through and fails the second."""
    fa, a = source(64, 48, 3, seed=5)
    ra = core.lgcr.Recon(a)
    assert same(ra.get_frame(0), recon_ref(64, 48, fa[0]))
    monkeypatch.setenv("LGCR_INJECT_FAULT", "launch:2")
    assert same(ra.get_frame(1), recon_ref(64, 48, fa[1]))
    with pytest.raises(Error, match="injected"):
        ra.get_frame(2)
    assert same(ra.get_frame(2), recon_ref(64, 48, fa[2]))


@pytest.mark.parametrize("kind", ["TRecon", "Sharpen", "Downsample"])
def test_launch_fault_other_filters(device, monkeypatch, kind):
    prefix = {"TRecon": "TRecon", "Sharpen": "LGCR Sharpen", "Downsample": "Downsample"}[kind]
    frames, clip = source(64, 48, 3, seed=6, ss=(0, 0) if kind == "Downsample" else (1, 1))
    filt = getattr(core.lgcr, kind)(clip)
    first = [np.asarray(p).copy() for p in filt.get_frame(1)]
    monkeypatch.setenv("LGCR_INJECT_FAULT", "launch:1")
    with pytest.raises(Error, match=rf"^{prefix}: .*injected"):
        filt.get_frame(1)
    again = filt.get_frame(1)
    assert all(np.array_equal(np.asarray(again[i]), first[i]) for i in range(3))


def test_alloc_fault_at_creation(device, monkeypatch):
    """A clip's first workspace is allocated when the filter is created: a
    failed allocation is a creation error, and the next creation works."""
    fa, a = source(80, 48, 2, seed=7)
    monkeypatch.setenv("LGCR_INJECT_FAULT", "alloc:1")
    with pytest.raises(Error, match=r"^LGCR: .*hipMalloc"):
        core.lgcr.Recon(a)
    ra = core.lgcr.Recon(a)
    assert same(ra.get_frame(0), recon_ref(80, 48, fa[0]))
    assert same(ra.get_frame(1), recon_ref(80, 48, fa[1]))


@pytest.mark.parametrize("n", [1, 2, 3])
def test_alloc_fault_in_serving(device, monkeypatch, n):
    """A failed workspace allocation while serving (an admission or a clip's
    extra workspace) delays frames but serves every one, identical to a run
    without the fault."""
    specs = serve.make_mix(3, 4, seed=13, resolutions=[(256, 144), (320, 180)], kinds=("recon",))
    reqs = serve.poisson_requests(specs, 2000.0, seed=13)
    clean = {}
    clips = [serve.ResidentClip(device, s, max_inflight=2, threads=4) for s in specs]
    assert serve.serve(clips, reqs, mode="continuous", keep=clean)["frames"] == len(reqs)
    for c in clips:
        c.release()
    monkeypatch.setenv("LGCR_INJECT_FAULT", f"alloc:{n}")
    got = {}
    st = serve.serve(clips, reqs, mode="continuous", keep=got)
    assert st["frames"] == len(reqs) and st["failed"] == 0
    for key, planes in got.items():
        assert all(np.array_equal(x, y) for x, y in zip(planes, clean[key])), key


def test_serving_skips_failed_frame(device, monkeypatch):
    """In continuous serving a failed submission is reported and the loop
    serves every other frame, identical to a run without the fault."""
    specs = serve.make_mix(3, 4, seed=11, resolutions=[(256, 144), (320, 180)], kinds=("recon",))
    reqs = serve.poisson_requests(specs, 600.0, seed=11)
    clips = [serve.ResidentClip(device, s, max_inflight=2, threads=4) for s in specs]
    clean = {}
    st = serve.serve(clips, reqs, mode="continuous", keep=clean)
    assert st["frames"] == len(reqs) and st["failed"] == 0
    monkeypatch.setenv("LGCR_INJECT_FAULT", "launch:5")
    got, failures = {}, []
    st = serve.serve(clips, reqs, mode="continuous", keep=got, failures=failures)
    assert st["failed"] == 1 and st["frames"] == len(reqs) - 1
    r, msg = failures[0]
    assert "injected" in msg
    assert set(got) == set(clean) - {(r.clip, r.frame)}
    for key, planes in got.items():
        assert all(np.array_equal(x, y) for x, y in zip(planes, clean[key])), key
