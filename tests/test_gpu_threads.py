"""Concurrent host-plane frames on one clip: eight threads share one
Downsample clip and one Sharpen clip created with eight workspaces (the
clip starts with one and adds the others while frames are in flight).
Every frame must equal the oracle's."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from lgcr_amd import abi, synth

pytestmark = pytest.mark.gpu

THREADS = 8
FRAMES = 24


def frames444(w, h, n):
    return [synth.clip_frame(w, h, ss_w=0, ss_h=0, bits=10, seed=100 + k, polygons=12) for k in range(n)]


def test_downsample_eight_threads(device):
    w, h = 192, 128
    frames = frames444(w, h, FRAMES)
    params = dict(quality=2, kernel="spline36", strength=1.0, loc="left")
    clip = abi.DownsampleClip(device, abi.downsample_params(**params), abi.make_format(np.uint16, w, h, 0, 0, 10),
                              max_inflight=THREADS)
    try:
        with ThreadPoolExecutor(THREADS) as ex:
            got = list(ex.map(lambda f: clip.downsample_host(*f), frames))
    finally:
        clip.close()
    for k, f in enumerate(frames):
        ref = oracle.downsample_frame(oracle.downsample_params(**params), *f, bits=10)
        for i in range(3):
            assert np.array_equal(ref[i], got[k][i]), (k, "YUV"[i])


def test_sharpen_eight_threads(device):
    w, h = 192, 128
    frames = [synth.clip_frame(w, h, bits=8, seed=200 + k, polygons=12) for k in range(FRAMES)]
    clip = abi.SharpenClip(device, abi.sharpen_params(), abi.make_format(np.uint8, w, h), max_inflight=THREADS)
    try:
        with ThreadPoolExecutor(THREADS) as ex:
            got = list(ex.map(lambda f: clip.sharpen_host(*f), frames))
    finally:
        clip.close()
    for k, f in enumerate(frames):
        ref = oracle.sharpen_frame(oracle.sharpen_params(), *f)
        for i in range(3):
            assert np.array_equal(ref[i], got[k][i]), (k, "YUV"[i])
