"""GPU parity with intermediate capture off (the production and benchmark
configuration).  Without capture the backend may skip work that only the
stage fetches observe, e.g. the compacted index list when the tiled guided
pass reads the trust mask bits directly; these cases check the frames and
the workset statistics against the oracle on that path."""
import numpy as np
import pytest

import oracle
from lgcr_amd import abi
from parity import compare_planes, frame, run_both

pytestmark = pytest.mark.gpu

CASES = [
    # (id, width, height, ss_w, ss_h, bits, params)
    ("p8-420", 96, 64, 1, 1, 8, {}),
    ("p10-420-bp", 160, 96, 1, 1, 10, {"bp": 0.5}),
    ("p16-420", 96, 64, 1, 1, 16, {}),
    ("p8-422", 96, 64, 1, 0, 8, {}),
    ("p8-444", 96, 64, 0, 0, 8, {}),
    ("ps-420", 96, 64, 1, 1, 32, {}),
    ("p8-420-odd", 98, 46, 1, 1, 8, {}),
    ("p8-420-tiny", 2, 2, 1, 1, 8, {}),
    ("p8-420-center", 96, 64, 1, 1, 8, {"shift_x": 0.0}),
    ("p8-420-top", 96, 64, 1, 1, 8, {"shift_y": -0.5}),
    ("p8-420-spline36", 96, 64, 1, 1, 8, {"kernel": "spline36"}),
    ("p8-420-noms", 96, 64, 1, 1, 8, {"ms": 0.0}),
    ("p8-420-algo4", 96, 64, 1, 1, 8, {"algo": 4}),
    ("ps-420-algo4", 96, 64, 1, 1, 32, {"algo": 4}),
    ("p8-420-algo6", 96, 64, 1, 1, 8, {"algo": 6}),
    ("p8-420-upscale", 96, 64, 1, 1, 8, {"width": 160, "height": 120}),
    ("p8-420-bm", 96, 64, 1, 1, 8, {"bm": 1, "cn": 3.0}),
    ("p8-420-sigma", 96, 64, 1, 1, 8, {"sigma": 0.002}),
    # back-projection inside the guided tiles: every sample type, a frame
    # whose width and height are not multiples of the 8x8 tile, and 4:2:2
    # where bp does not apply (packing only)
    ("p8-420-bp", 96, 64, 1, 1, 8, {"bp": 1.0}),
    ("ps-420-bp", 96, 64, 1, 1, 32, {"bp": 0.5}),
    ("p16-420-bp-ragged", 100, 58, 1, 1, 16, {"bp": 0.7}),
    # algo 4 with integer output: the selector is final in the tiles, which
    # back-project and pack the output themselves
    ("p10-420-algo4-bp", 160, 96, 1, 1, 10, {"algo": 4, "bp": 0.5}),
    ("p8-420-algo4-bp-ragged", 100, 58, 1, 1, 8, {"algo": 4, "bp": 0.5}),
    ("p16-422-algo4", 96, 64, 1, 0, 16, {"algo": 4}),
    ("p8-422-bp", 96, 64, 1, 0, 8, {"bp": 0.5}),
]


@pytest.fixture
def no_capture(device, monkeypatch):
    monkeypatch.setenv("LGCR_CAPTURE", "0")
    return device


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_recon_no_capture(no_capture, case):
    tag, w, h, ssw, ssh, bits, params = case
    params = dict(params)
    cn = params.pop("cn", 0.0)
    y, u, v = frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=11, polygons=12, chroma_noise=cn,
                    luma_noise=cn if cn else None)
    o_out, dbg, g_out, clip = run_both(y, u, v, ss_w=ssw, ss_h=ssh, bits=None if bits in (8, 16, 32) else bits,
                                       device=no_capture, **params)
    try:
        for i, name in enumerate("YUV"):
            n, mx = compare_planes(o_out[i], g_out[i], bits == 32)
            print(f"{tag} {name}: mismatches={n} max_abs={mx:g}")
            assert n == 0
        # workset statistics of the same frame
        _, st = clip.recon_host(y, u, v, stats=True)
        if "indices" in dbg and params.get("sparse", 1):
            print(f"{tag}: active outputs {st.active_output_pixels} (oracle {len(dbg['indices'])}), "
                  f"active chroma {st.active_chroma_pixels} (oracle {int(dbg['chroma_mask'].sum())})")
            assert st.active_output_pixels == len(dbg["indices"])
            assert st.active_chroma_pixels == int(dbg["chroma_mask"].sum())
    finally:
        clip.close()


def test_bench_frame_no_capture(no_capture):
    """The benchmark's frame: 1920x1080 YUV420P10, algo 2, bp 0.5."""
    y, u, v = frame(1920, 1080, bits=10, seed=3, polygons=40)
    o_out, dbg, g_out, clip = run_both(y, u, v, bits=10, device=no_capture, bp=0.5)
    try:
        for i, name in enumerate("YUV"):
            n, mx = compare_planes(o_out[i], g_out[i], False)
            print(f"1080p {name}: mismatches={n} max_abs={mx:g}")
            assert n == 0
        _, st = clip.recon_host(y, u, v, stats=True)
        assert st.active_output_pixels == len(dbg["indices"])
        assert st.active_chroma_pixels == int(dbg["chroma_mask"].sum())
    finally:
        clip.close()


def test_algo4_frame_no_capture(no_capture):
    """1920x1080 YUV420P10, algo 4, bp 0.5 (the algo-4 benchmark workload):
    selector, back-projection and packing in the tiles, bit-exact."""
    y, u, v = frame(1920, 1080, bits=10, seed=3, polygons=40)
    o_out, _, g_out, clip = run_both(y, u, v, bits=10, device=no_capture, algo=4, bp=0.5)
    try:
        for i, name in enumerate("YUV"):
            d = np.abs(o_out[i].astype(np.int64) - g_out[i].astype(np.int64))
            print(f"1080p algo 4 {name}: max_abs={d.max()}")
            assert d.max() == 0
    finally:
        clip.close()


TRECON_FORMATS = {"420p8": (1, 1, 8), "420p10": (1, 1, 10), "422p16": (1, 0, 16), "444ps": (0, 0, 32)}
# sparse 1: plain base in the tiles; sparse 0 (the TRecon default): the dense
# temporal tiles write the packed output themselves
TRECON_CASES = [("420p10", 1, 96, 64), ("444ps", 1, 96, 64), ("420p8", 0, 96, 64), ("420p10", 0, 96, 64),
                ("422p16", 0, 98, 62), ("444ps", 0, 96, 64)]


@pytest.mark.parametrize("fmt,sparse,w,h", TRECON_CASES, ids=[f"{f}-sparse{s}-{w}x{h}" for f, s, w, h in TRECON_CASES])
def test_trecon_no_capture(no_capture, fmt, sparse, w, h):
    ssw, ssh, bits = TRECON_FORMATS[fmt]
    frames = [frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=5, frame=k, motion=True) for k in range(3)]
    tp = abi.trecon_params(sparse=sparse)
    clip = abi.Clip(no_capture, tp, abi.make_format(frames[0][0].dtype, w, h, ssw, ssh,
                                                    None if bits in (8, 16, 32) else bits), temporal=True)
    try:
        got = clip.trecon_host(frames[1], [frames[0], frames[2]])
        ref = oracle.trecon_frame(oracle.trecon_params(sparse=sparse), frames[1], [frames[0], frames[2]],
                                  bits=None if bits in (8, 16, 32) else bits, ss_w=ssw, ss_h=ssh)
        for i in range(3):
            n, mx = compare_planes(ref[i], got[i], bits == 32)
            print(f"trecon {fmt} sparse {sparse} {w}x{h} plane {i}: mismatches={n} max_abs={mx:g}")
            assert n == 0
    finally:
        clip.close()
