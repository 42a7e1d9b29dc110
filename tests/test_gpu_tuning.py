"""Tile-shape tuning: every candidate shape of the tunable kernels (Sobel
tensor 64x{8,16,32}, mutual gate 16x16 / 32x8 / 64x4) gives the oracle's
frames, the autotuner picks one of them and remembers it per shape key."""
import itertools

import numpy as np
import pytest

import oracle
from lgcr_amd import abi
from parity import compare_planes, frame

pytestmark = pytest.mark.gpu

SHAPES = list(itertools.product((8, 16, 32), (16, 32, 64)))
CASES = [
    # (id, width, height, ss_w, ss_h, bits, params)
    ("p8-420-odd", 98, 46, 1, 1, 8, {}),
    ("p10-422", 160, 96, 1, 0, 10, {"bp": 0.5}),
    ("ps-444-algo4", 96, 64, 0, 0, 32, {"algo": 4}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_every_shape_matches_oracle(device, case):
    tag, w, h, ssw, ssh, bits, params = case
    y, u, v = frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=31, polygons=10)
    b = None if bits in (8, 16, 32) else bits
    ref = oracle.recon_frame(oracle.make_params(w, h, **params), y, u, v, bits=b, ss_w=ssw, ss_h=ssh)
    clip = abi.Clip(device, abi.default_params(**params), abi.make_format(y.dtype, w, h, ssw, ssh, b))
    try:
        for rows, cols in SHAPES:
            clip.set_tuning(rows, cols)
            assert clip.tuning() == (rows, cols)
            got = clip.recon_host(y, u, v)
            for i in range(3):
                cnt, mx = compare_planes(ref[i], got[i], bits == 32)
                assert cnt == 0, f"{tag} shape {rows}/{cols} plane {i}: {cnt} mismatches (max {mx:g})"
    finally:
        clip.close()


def test_trecon_shapes_match_oracle(device):
    frames = [frame(96, 64, bits=10, seed=8, frame=k, motion=True) for k in range(3)]
    ref = oracle.trecon_frame(oracle.trecon_params(), frames[1], [frames[0], frames[2]], bits=10)
    clip = abi.Clip(device, abi.trecon_params(), abi.make_format(np.uint16, 96, 64, bits=10), temporal=True)
    try:
        for rows, cols in ((8, 64), (32, 16)):
            clip.set_tuning(rows, cols)
            got = clip.trecon_host(frames[1], [frames[0], frames[2]])
            for i in range(3):
                cnt, _ = compare_planes(ref[i], got[i], False)
                assert cnt == 0, f"shape {rows}/{cols} plane {i}"
    finally:
        clip.close()


def test_invalid_shape_is_rejected(device):
    clip = abi.Clip(device, abi.default_params(), abi.make_format(np.uint8, 64, 32))
    try:
        with pytest.raises(abi.LgcrError):
            clip.set_tuning(5, 32)
        with pytest.raises(abi.LgcrError):
            clip.set_tuning(8, 48)
        assert clip.tuning() == (16, 32)
    finally:
        clip.close()


def test_autotune_picks_a_candidate_and_remembers_it(device):
    w, h = 1280, 720
    y, u, v = frame(w, h, bits=10, seed=2, polygons=30)
    fmt = abi.make_format(np.uint16, w, h, bits=10)
    clip = abi.Clip(device, abi.default_params(bp=0.5), fmt)
    try:
        shape = clip.autotune_host(y, u, v)
        print("autotuned 1280x720 P10:", shape)
        assert shape in SHAPES
        ref = oracle.recon_frame(oracle.make_params(w, h, bp=0.5), y, u, v, bits=10)
        got = clip.recon_host(y, u, v)
        for i in range(3):
            assert compare_planes(ref[i], got[i], False)[0] == 0
    finally:
        clip.close()
    again = abi.Clip(device, abi.default_params(bp=0.5), fmt)
    try:
        assert again.tuning() == shape
    finally:
        again.close()
