"""GPU parity: the HIP backend (through the C ABI) against the oracle.

Bars: float planes within 1e-5 absolute, integer planes within 1 code
value, masks / workset indices identical.  Every mismatch is counted and
printed.
"""
import threading

import numpy as np
import pytest

from parity import compare_planes, frame, run_both

pytestmark = pytest.mark.gpu


def _report(tag, o_out, g_out, is_float):
    worst = 0
    for i, name in enumerate("YUV"):
        n, mx = compare_planes(o_out[i], g_out[i], is_float)
        print(f"{tag} {name}: mismatches={n} max_abs={mx:g}")
        worst = max(worst, n)
    return worst


def test_config1_stagewise(device):
    """640x360 YUV420P8, defaults (algo 2, lanczos 3, strength 0.8, sparse 1)."""
    y, u, v = frame(640, 360, bits=8, seed=0)
    o_out, dbg, g_out, clip = run_both(y, u, v, device=device)
    try:
        for name in ("y", "u", "v", "jxx", "jxy", "jyy", "lc"):
            g = clip.fetch(name)
            n, mx = compare_planes(dbg[name], g, True)
            print(f"stage {name}: mismatches={n} max_abs={mx:g}")
            assert n == 0, name
        for name in ("seed", "mask", "chroma_mask"):
            g = clip.fetch(name, np.uint8)
            diff = int((g != dbg[name]).sum())
            print(f"stage {name}: differing={diff}")
            assert diff == 0, name
        idx = clip.fetch("indices", np.uint32)
        assert idx.shape == dbg["indices"].shape and np.array_equal(idx, dbg["indices"])
        ms = clip.fetch("ms")
        n, mx = compare_planes(dbg["ms"], ms, True)
        print(f"stage ms: mismatches={n} max_abs={mx:g}")
        assert n == 0
        for name in ("plain_u", "plain_v", "out_u", "out_v"):
            n, mx = compare_planes(dbg[name], clip.fetch(name), True)
            print(f"stage {name}: mismatches={n} max_abs={mx:g}")
            assert n == 0, name
        assert _report("config1", o_out, g_out, False) == 0
        assert np.array_equal(g_out[0], y)
    finally:
        clip.close()


CASES = [
    # (id, width, height, ss_w, ss_h, bits, params)
    ("p8-420-default", 96, 64, 1, 1, 8, {}),
    ("p10-420-default", 96, 64, 1, 1, 10, {}),
    ("p16-420-default", 96, 64, 1, 1, 16, {}),
    ("p8-422-default", 96, 64, 1, 0, 8, {}),
    ("p8-444-default", 96, 64, 0, 0, 8, {}),
    ("ps-420-default", 96, 64, 1, 1, 32, {}),
    ("p8-420-strength0", 96, 64, 1, 1, 8, {"strength": 0.0}),
    ("p10-420-bp", 96, 64, 1, 1, 10, {"bp": 0.5}),
    ("ps-420-bp1", 96, 64, 1, 1, 32, {"bp": 1.0}),
    ("p8-420-dense", 96, 64, 1, 1, 8, {"sparse": 0}),
    ("p8-420-noms", 96, 64, 1, 1, 8, {"ms": 0.0}),
    ("p8-420-noridge-cedge", 96, 64, 1, 1, 8, {"ridge": 0, "cedge": 1}),
    ("p8-420-center", 96, 64, 1, 1, 8, {"shift_x": 0.0}),
    ("p8-420-top", 96, 64, 1, 1, 8, {"shift_y": -0.5}),
    ("p8-420-bicubic", 96, 64, 1, 1, 8, {"kernel": "bicubic"}),
    ("p8-420-spline16", 96, 64, 1, 1, 8, {"kernel": "spline16"}),
    ("p8-420-spline36", 96, 64, 1, 1, 8, {"kernel": "spline36"}),
    ("p8-420-bilinear", 96, 64, 1, 1, 8, {"kernel": "bilinear"}),
    ("p8-420-lanczos4", 96, 64, 1, 1, 8, {"kernel": "lanczos", "taps": 4}),
    ("p8-420-lanczos1", 96, 64, 1, 1, 8, {"kernel": "lanczos", "taps": 1}),
    ("p8-420-ar-off", 96, 64, 1, 1, 8, {"ar": -1.0}),
    ("p8-420-ar-margin", 96, 64, 1, 1, 8, {"ar": 0.02}),
    ("p8-420-tiny", 2, 2, 1, 1, 8, {}),
    ("p8-444-2x64", 2, 64, 0, 0, 8, {}),
    ("p8-444-64x2", 64, 2, 0, 0, 8, {}),
    ("p8-420-odd-rows", 98, 46, 1, 1, 8, {}),
    ("p8-420-upscale", 96, 64, 1, 1, 8, {"width": 160, "height": 120}),
    ("p8-420-downscale", 96, 64, 1, 1, 8, {"width": 64, "height": 40}),
    ("ps-420-upscale-l4", 96, 64, 1, 1, 32, {"width": 192, "height": 128, "kernel": "lanczos", "taps": 4}),
    ("p10-422-upscale-bicubic", 96, 64, 1, 0, 10, {"width": 150, "height": 100, "kernel": "bicubic"}),
    # radial (jinc) base and guided weights
    ("p8-420-jinc", 96, 64, 1, 1, 8, {"kernel": "jinc"}),
    ("ps-420-jinc-strength0", 96, 64, 1, 1, 32, {"kernel": "jinc", "strength": 0.0}),
    ("p8-420-jinc4-dense", 96, 64, 1, 1, 8, {"kernel": "jinc", "taps": 4, "sparse": 0}),
    ("p8-420-jinc-upscale", 96, 64, 1, 1, 8, {"kernel": "jinc", "width": 192, "height": 128}),
    ("ps-444-jinc-downscale", 96, 64, 0, 0, 32, {"kernel": "jinc", "width": 60, "height": 40}),
    # algo 4 selector: integer (single expression) and float (metadata + blend)
    ("p8-420-algo4", 96, 64, 1, 1, 8, {"algo": 4}),
    ("p10-420-algo4-bp", 96, 64, 1, 1, 10, {"algo": 4, "bp": 0.5}),
    ("p8-420-algo4-cedge", 96, 64, 1, 1, 8, {"algo": 4, "cedge": 1}),
    ("ps-420-algo4", 96, 64, 1, 1, 32, {"algo": 4}),
    ("ps-420-algo4-dense", 96, 64, 1, 1, 32, {"algo": 4, "sparse": 0}),
    ("ps-420-algo4-cedge", 96, 64, 1, 1, 32, {"algo": 4, "cedge": 1}),
    ("ps-420-algo4-compressed", 192, 96, 1, 1, 32, {"algo": 4, "poly": 2}),
    ("p8-420-algo4-compressed", 192, 96, 1, 1, 8, {"algo": 4, "poly": 2}),
    ("ps-422-algo4-jinc", 96, 64, 1, 0, 32, {"algo": 4, "kernel": "jinc"}),
    ("p8-420-algo4-upscale", 96, 64, 1, 1, 8, {"algo": 4, "width": 144, "height": 96}),
    # algo 6 detail transfer: dense and sparse affine ROI, resize via source grid
    ("p8-420-algo6", 96, 64, 1, 1, 8, {"algo": 6}),
    ("p10-420-algo6-bp", 96, 64, 1, 1, 10, {"algo": 6, "bp": 0.5}),
    ("ps-420-algo6", 96, 64, 1, 1, 32, {"algo": 6}),
    ("ps-420-algo6-dense", 96, 64, 1, 1, 32, {"algo": 6, "sparse": 0}),
    ("ps-420-algo6-sparse-roi", 192, 96, 1, 1, 32, {"algo": 6, "poly": 2}),
    ("p8-420-algo6-bilinear", 96, 64, 1, 1, 8, {"algo": 6, "kernel": "bilinear"}),
    ("p8-420-algo6-jinc", 96, 64, 1, 1, 8, {"algo": 6, "kernel": "jinc"}),
    ("p8-422-algo6", 96, 64, 1, 0, 8, {"algo": 6}),
    ("ps-444-algo6", 96, 64, 0, 0, 32, {"algo": 6}),
    ("ps-420-algo6-qgate", 96, 64, 1, 1, 32, {"algo": 6, "qgate": 0.3, "ms": 0.5}),
    ("ps-420-resize-jinc3-algo6", 120, 68, 1, 1, 32, {"algo": 6, "kernel": "jinc", "width": 240, "height": 136}),
    ("ps-420-resize-lanczos4-algo6", 120, 68, 1, 1, 32,
     {"algo": 6, "kernel": "lanczos", "taps": 4, "width": 240, "height": 136}),
    ("p8-420-downscale-algo6", 96, 64, 1, 1, 8, {"algo": 6, "width": 72, "height": 48}),
    # bm=True collaborative chroma filter after each algo, before bp
    ("p8-420-bm", 96, 64, 1, 1, 8, {"bm": 1, "cn": 3.0}),
    ("p10-420-bm-bp", 96, 64, 1, 1, 10, {"bm": 1, "bp": 0.5, "cn": 6.0}),
    ("ps-444-bm-strength0", 64, 64, 0, 0, 32, {"bm": 1, "strength": 0.0, "cn": 0.006}),
    ("p8-422-bm-dense", 96, 64, 1, 0, 8, {"bm": 1, "sparse": 0, "cn": 3.0}),
    ("p8-420-bm-odd", 98, 46, 1, 1, 8, {"bm": 1, "cn": 3.0}),
    ("p8-444-bm-narrow", 6, 64, 0, 0, 8, {"bm": 1, "cn": 3.0}),
    ("p8-420-bm-upscale", 96, 64, 1, 1, 8, {"bm": 1, "width": 160, "height": 120, "cn": 3.0}),
    ("ps-420-bm-downscale-jinc", 96, 64, 1, 1, 32, {"bm": 1, "kernel": "jinc", "width": 72, "height": 40,
                                                   "cn": 0.01}),
    ("ps-420-bm-algo4", 96, 64, 1, 1, 32, {"bm": 1, "algo": 4, "cn": 0.01}),
    ("p8-420-bm-algo6", 96, 64, 1, 1, 8, {"bm": 1, "algo": 6, "cn": 3.0}),
    ("ps-420-bm-sigma", 96, 64, 1, 1, 32, {"bm": 1, "sigma": 0.002, "cn": 0.004}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_recon_matrix(device, case):
    tag, w, h, ssw, ssh, bits, params = case
    params = dict(params)
    poly = params.pop("poly", 12)
    cn = params.pop("cn", 0.0)
    y, u, v = frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=7, polygons=poly, chroma_noise=cn,
                    luma_noise=cn if cn else None)
    o_out, _, g_out, clip = run_both(y, u, v, ss_w=ssw, ss_h=ssh, bits=None if bits in (8, 16, 32) else bits,
                                     device=device, **params)
    clip.close()
    assert _report(tag, o_out, g_out, bits == 32) == 0


def test_config2_frame_1080p_p10(device):
    """One full 1920x1080 YUV420P10 frame, algo 2 with bp 0.5."""
    y, u, v = frame(1920, 1080, bits=10, seed=0)
    o_out, dbg, g_out, clip = run_both(y, u, v, bits=10, device=device, bp=0.5)
    try:
        idx = clip.fetch("indices", np.uint32)
        assert np.array_equal(idx, dbg["indices"])
        assert np.array_equal(clip.fetch("mask", np.uint8), dbg["mask"])
        n_ou, mx = compare_planes(dbg["out_u"], clip.fetch("out_u"), True)
        print(f"1080p out_u float mismatches={n_ou} max_abs={mx:g}")
        assert n_ou == 0
        assert _report("1080p-p10", o_out, g_out, False) == 0
    finally:
        clip.close()


@pytest.mark.parametrize("kernel,taps", [("jinc", 3), ("lanczos", 4)])
def test_config3_resize_algo6_float(device, kernel, taps):
    """1920x1080 YUV420PS (noise sigma 0.002) -> 3840x2160 YUV444PS with
    algo 6: every output plane within 1e-5 of the oracle (which takes about
    10 s for the jinc case on one host core)."""
    y, u, v = frame(1920, 1080, bits=32, seed=2)
    o_out, _, g_out, clip = run_both(y, u, v, device=device, algo=6, kernel=kernel, taps=taps, width=3840,
                                     height=2160)
    clip.close()
    assert g_out[1].shape == (2160, 3840)
    assert _report(f"config3-{kernel}{taps}-algo6", o_out, g_out, True) == 0


@pytest.mark.parametrize("algo", [2, 4, 6])
def test_config2_algo_1080p_p10_float_stage(device, algo):
    """1080p YUV420P10 with bp 0.5 for algo 2 and 4: float planes before
    packing within 1e-5, packed planes within 1 code."""
    y, u, v = frame(1920, 1080, bits=10, seed=1, frame=3)
    o_out, dbg, g_out, clip = run_both(y, u, v, bits=10, device=device, bp=0.5, algo=algo)
    try:
        for name in ("out_u", "out_v"):
            n, mx = compare_planes(dbg[name], clip.fetch(name), True)
            print(f"1080p algo{algo} {name}: mismatches={n} max_abs={mx:g}")
            assert n == 0
        assert _report(f"1080p-algo{algo}", o_out, g_out, False) == 0
    finally:
        clip.close()


def test_config4_bm_1080p_422p8(device):
    """1920x1080 YUV422P8 with 3-LSB noise on every plane, algo 2, bm=True,
    bp 0.5: the float planes after the collaborative filter within 1e-5 and
    the packed planes within 1 code."""
    y, u, v = frame(1920, 1080, ss_w=1, ss_h=0, bits=8, seed=3, luma_noise=3.0, chroma_noise=3.0)
    o_out, dbg, g_out, clip = run_both(y, u, v, ss_w=1, ss_h=0, device=device, bm=1, bp=0.5)
    try:
        for name in ("out_u", "out_v"):
            n, mx = compare_planes(dbg[name], clip.fetch(name), True)
            print(f"1080p422 bm {name}: mismatches={n} max_abs={mx:g}")
            assert n == 0
        assert _report("1080p422-bm", o_out, g_out, False) == 0
    finally:
        clip.close()


def test_concurrent_calls_identical(device):
    """fmParallel: concurrent calls on one clip give byte-identical frames
    (test/test_concurrency.py:44-55)."""
    from lgcr_amd import abi
    frames = [frame(160, 96, bits=8, seed=3, frame=k) for k in range(6)]
    clip = abi.Clip(device, abi.default_params(), abi.make_format(np.uint8, 160, 96))
    try:
        ref = [clip.recon_host(*f) for f in frames]
        results = [None] * 24
        errors = []

        def work(i):
            try:
                results[i] = clip.recon_host(*frames[i % 6])
            except Exception as e:  # pragma: no cover
                errors.append(e)

        ts = [threading.Thread(target=work, args=(i,)) for i in range(24)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errors
        for i in range(24):
            for p in range(3):
                assert np.array_equal(results[i][p], ref[i % 6][p])
    finally:
        clip.close()


def test_batch_device_matches_host(device):
    """lgcr_recon_batch on HBM-resident frames == per-frame host path."""
    import torch
    from lgcr_amd import abi
    w, h = 320, 180
    frames = [frame(w, h, bits=10, seed=5, frame=k) for k in range(5)]
    clip = abi.Clip(device, abi.default_params(bp=0.5), abi.make_format(np.uint16, w, h, bits=10))
    try:
        ref = [clip.recon_host(*f) for f in frames]
        dev_in, dev_out, fi, fo = [], [], [], []
        for f in frames:
            planes = [torch.from_numpy(p.astype(np.int32)).to(torch.int16).cuda() for p in f]
            outs = [torch.empty((h, w), dtype=torch.int16, device="cuda") for _ in range(2)]
            dev_in.append(planes)
            dev_out.append(outs)
            fi.append([(t.data_ptr(), t.stride(0) * 2) for t in planes])
            fo.append([(0, 0)] + [(t.data_ptr(), t.stride(0) * 2) for t in outs])
        clip.recon_batch(fi, fo)
        torch.cuda.synchronize()
        for k in range(5):
            for p in range(2):
                got = dev_out[k][p].cpu().numpy().view(np.uint16)
                assert np.array_equal(got, ref[k][p + 1])
    finally:
        clip.close()
