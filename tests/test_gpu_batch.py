"""Multi-frame launches: lgcr_recon_batch with batching on runs the fused
tile pass (plain base, guided correction, back-projection, packing) of a
group of frames in ONE launch (one wave per tile over a (tiles, frames)
grid, or a persistent grid whose waves stride over the tiles of every
frame), and on that path the earlier stages (conversion, Sobel tensor,
footprint luma, trust mask, mutual gate) as one launch per stage for up to
eight frames.  Every frame must equal the frame-by-frame launches
bit for bit, and the oracle; groups of 1..8 frames, a batch that does not
fill its last group, ragged frames, every fused output type, and frames
whose guided pass is not the fused tile pass (those complete on their own
streams) are covered."""
import numpy as np
import pytest
import torch

import oracle
from lgcr_amd import abi
from parity import frame

pytestmark = pytest.mark.gpu

CASES = [
    # (id, width, height, ss_w, ss_h, bits, params, max_inflight, frames, batching)
    ("p10-420-bp-g4", 320, 180, 1, 1, 10, {"bp": 0.5}, 8, 9, 8),
    ("p10-420-g1", 320, 180, 1, 1, 10, {}, 2, 3, 1),
    ("p8-420-ragged-g3", 198, 118, 1, 1, 8, {"bp": 0.5}, 6, 7, 3),
    ("p16-420-g2", 160, 96, 1, 1, 16, {"bp": 0.5}, 4, 5, 2),
    ("ps-420-g4", 160, 96, 1, 1, 32, {"bp": 0.5}, 8, 6, 4),
    ("p10-422-g2", 160, 96, 1, 0, 10, {}, 4, 4, 2),
    ("p10-420-dense-g2", 160, 96, 1, 1, 10, {"sparse": 0}, 4, 4, 2),
    # groups past eight frames: the earlier stages go in launches of eight
    ("p10-420-bp-g12", 160, 96, 1, 1, 10, {"bp": 0.5}, 12, 14, 12),
    # no mutual gate: the group launches stop at the trust mask
    ("p8-420-ms0-g3", 198, 118, 1, 1, 8, {"ms": 0.0}, 6, 7, 3),
    # not the fused tile pass: algo 4 and bm finish on their own streams
    ("p10-420-algo4-g2", 160, 96, 1, 1, 10, {"algo": 4}, 4, 3, 2),
    ("p8-420-bm-g2", 160, 96, 1, 1, 8, {"bm": 1}, 4, 3, 2),
]


def _dev(a):
    t = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a)
    return t.to("cuda")


@pytest.fixture(params=["grid", "persistent"])
def no_capture(device, monkeypatch, request):
    """Stage capture off (the production path); both launch forms: one wave
    per tile over a (tiles, frames) grid, and the persistent grid."""
    monkeypatch.setenv("LGCR_CAPTURE", "0")
    monkeypatch.setenv("LGCR_BATCH_PERSISTENT", "1" if request.param == "persistent" else "0")
    return device


def _run(clip, din, n, ow, oh, bps, dt, same):
    tdt = {1: torch.uint8, 2: torch.int16, 4: torch.float32}[bps]
    outs = [[torch.full((oh, ow), 7, dtype=tdt, device="cuda") for _ in range(3)] for _ in range(n)]
    fin = [[(t.data_ptr(), t.stride(0) * bps) for t in f] for f in din]
    fout = [[(0, 0) if same else (o[0].data_ptr(), o[0].stride(0) * bps)] +
            [(t.data_ptr(), t.stride(0) * bps) for t in o[1:]] for o in outs]
    clip.recon_batch(fin, fout)
    torch.cuda.synchronize()
    return [[(t.cpu().numpy().view(dt) if bps == 2 else t.cpu().numpy()) for t in o] for o in outs]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_batched_launch_equals_frame_launches_and_oracle(no_capture, case):
    tag, w, h, ssw, ssh, bits, params, inflight, n, batching = case
    print(f"case {tag}: {w}x{h} bits {bits} frames {n} batching {batching}", flush=True)
    frames = [frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=31, frame=k, polygons=10) for k in range(n)]
    dt = frames[0][0].dtype
    bps = dt.itemsize
    fmt = abi.make_format(dt, w, h, ssw, ssh, None if bits in (8, 16, 32) else bits)
    clip = abi.Clip(no_capture, abi.default_params(**params), fmt, max_inflight=inflight)
    din = [[_dev(p) for p in f] for f in frames]
    try:
        direct = _run(clip, din, n, w, h, bps, dt, True)
        clip.set_batching(batching)
        assert clip.batching() == batching
        for rep in range(2):
            got = _run(clip, din, n, w, h, bps, dt, True)
            for k in range(n):
                for i in (1, 2):
                    assert np.array_equal(got[k][i], direct[k][i]), f"{tag} rep {rep} frame {k} plane {i}"
        for k in (0, n - 1):
            ref = oracle.recon_frame(oracle.make_params(w, h, **params), *frames[k],
                                     bits=None if bits in (8, 16, 32) else bits, ss_w=ssw, ss_h=ssh)
            for i in (1, 2):
                d = np.abs(ref[i].astype(np.float64) - got[k][i].astype(np.float64))
                print(f"{tag} frame {k} plane {i}: max_abs={d.max():g}")
                assert d.max() <= (1e-5 if bits == 32 else 0), f"{tag} frame {k} plane {i} vs oracle"
        clip.set_batching(0)
        again = _run(clip, din, n, w, h, bps, dt, True)
        for k in range(n):
            for i in (1, 2):
                assert np.array_equal(again[k][i], direct[k][i])
    finally:
        clip.close()


def test_batched_1080p_frames(no_capture):
    """Sixteen 1080p P10 frames (the benchmark frames) in groups of eight."""
    w, h, n = 1920, 1080, 16
    frames = [frame(w, h, bits=10, seed=0, frame=k) for k in range(n)]
    fmt = abi.make_format(np.uint16, w, h, bits=10)
    clip = abi.Clip(no_capture, abi.default_params(bp=0.5), fmt, max_inflight=16)
    din = [[_dev(p) for p in f] for f in frames]
    try:
        direct = _run(clip, din, n, w, h, 2, np.uint16, True)
        clip.set_batching(8)
        got = _run(clip, din, n, w, h, 2, np.uint16, True)
        for k in range(n):
            for i in (1, 2):
                assert np.array_equal(got[k][i], direct[k][i]), f"frame {k} plane {i}"
        ref = oracle.recon_frame(oracle.make_params(w, h, bp=0.5), *frames[5], bits=10)
        for i in (1, 2):
            assert np.array_equal(ref[i], got[5][i]), f"frame 5 plane {i} vs oracle"
    finally:
        clip.close()


def test_batching_argument_errors(no_capture):
    fmt = abi.make_format(np.uint16, 64, 32, bits=10)
    clip = abi.Clip(no_capture, abi.default_params(), fmt, max_inflight=2)
    try:
        with pytest.raises(abi.LgcrError, match="0..64"):
            clip.set_batching(65)
        with pytest.raises(abi.LgcrError):
            clip.set_batching(-1)
    finally:
        clip.close()
    tfmt = abi.make_format(np.uint16, 64, 32, bits=10)
    tclip = abi.Clip(no_capture, abi.trecon_params(), tfmt, temporal=True)
    try:
        with pytest.raises(abi.LgcrError, match="not a Recon clip"):
            tclip.set_batching(4)
    finally:
        tclip.close()
