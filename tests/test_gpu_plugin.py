"""The plugin surface (core.lgcr.Recon / TRecon / Sharpen) on the GPU, written
the way the reference's own scripts exercise its plugin: same-size luma
sharing and profile counts (test/test_plane_sharing.py), the safety
regressions (test/test_regressions.py), the collaborative stage switch
(test/test_bm.py), plus frame-property siting and oracle parity through the
plugin path."""
import numpy as np
import pytest

import oracle
from lgcr_amd import plugin
from lgcr_amd.plugin import FLOAT, INTEGER, YUV, Error, clip_from_planes, core, query_video_format
from parity import compare_planes

pytestmark = pytest.mark.gpu

F444 = query_video_format(YUV, FLOAT, 32, 0, 0)
F420 = query_video_format(YUV, FLOAT, 32, 1, 1)


def blank(fmt, width, height, length=1, color=(0.0, 0.0, 0.0), props=None):
    cw, ch = width >> fmt.subsampling_w, height >> fmt.subsampling_h
    planes = (np.full((height, width), color[0], fmt.dtype), np.full((ch, cw), color[1], fmt.dtype),
              np.full((ch, cw), color[2], fmt.dtype))
    return clip_from_planes([tuple(p.copy() for p in planes) for _ in range(length)], fmt, props)


def from_fill(fmt, width, height, length, fill, props=None):
    """A clip whose frame n has planes written by fill(n, [y, u, v])."""
    cw, ch = width >> fmt.subsampling_w, height >> fmt.subsampling_h
    frames = []
    for n in range(length):
        planes = [np.zeros((height, width), fmt.dtype), np.zeros((ch, cw), fmt.dtype),
                  np.zeros((ch, cw), fmt.dtype)]
        fill(n, planes)
        frames.append(tuple(planes))
    return clip_from_planes(frames, fmt, props)


def arrays(frame):
    return [np.asarray(frame[p]).copy() for p in range(3)]


def assert_finite(frame, label):
    for p in range(3):
        assert np.isfinite(np.asarray(frame[p])).all(), f"{label}: non-finite plane {p}"


# ---------------------------------------------------------------------------
# same-size luma sharing and profile counts (test/test_plane_sharing.py)
# ---------------------------------------------------------------------------

SHARING_FORMATS = [
    (INTEGER, 8, 1, 1, 34, 26, "420-8"),
    (INTEGER, 10, 1, 1, 34, 26, "420-10"),
    (INTEGER, 12, 1, 0, 34, 27, "422-12"),
    (INTEGER, 16, 0, 0, 35, 27, "444-16"),
    (FLOAT, 32, 1, 1, 34, 26, "420-float"),
    (FLOAT, 32, 1, 0, 34, 27, "422-float"),
    (FLOAT, 32, 0, 0, 35, 27, "444-float"),
]


@pytest.mark.parametrize("sample_type,bits,ssw,ssh,width,height,label", SHARING_FORMATS,
                         ids=[f[-1] for f in SHARING_FORMATS])
def test_plane_sharing(device, monkeypatch, sample_type, bits, ssw, ssh, width, height, label):
    monkeypatch.setenv("LGCR_PROFILE", "1")
    fmt = query_video_format(YUV, sample_type, bits, ssw, ssh)

    def fill(n, planes):
        for p, values in enumerate(planes):
            yy, xx = np.indices(values.shape)
            if sample_type == FLOAT:
                values[:] = (0.1 * p + xx * 0.003 + yy * 0.005).astype(np.float32)
            else:
                peak = (1 << bits) - 1
                values[:] = ((xx * 17 + yy * 29 + p * 101) % peak).astype(values.dtype)

    source = from_fill(fmt, width, height, 1, fill, props={"_ChromaLocation": 0})
    source_y = source.get_frame(0)[0]
    for algo in (2, 4, 6):
        result = core.lgcr.Recon(source, width=width, height=height, algo=algo, sparse=1).get_frame(0)
        result_y = np.asarray(result[0])
        assert result_y.__array_interface__["data"][0] == source_y.__array_interface__["data"][0], \
            f"luma plane not shared: {label}/algo={algo}"
        assert result_y.tobytes() == source_y.tobytes()
        assert "_ChromaLocation" not in result.props, f"stale chroma location: {label}/algo={algo}"
        expected_input = (0 if sample_type == FLOAT else
                          width * height + 2 * ((width + (1 << ssw) - 1) >> ssw) * ((height + (1 << ssh) - 1) >> ssh))
        expected_output = 0 if sample_type == FLOAT else 2 * width * height
        assert result.props["_LGCR_input_conversion_pixels"] == expected_input
        assert result.props["_LGCR_output_conversion_pixels"] == expected_output
        assert result.props["_LGCR_luma_resample_pixels"] == 0
        assert result.props["_LGCR_total_us"] > 0.0
        assert 0.0 <= result.props["_LGCR_sparse_active_ratio"] <= 1.0


def test_profile_props_absent_without_env(device, monkeypatch):
    monkeypatch.delenv("LGCR_PROFILE", raising=False)
    out = core.lgcr.Recon(blank(F420, 32, 16, color=(0.4, 0.1, -0.1))).get_frame(0)
    assert not any(k.startswith("_LGCR_") for k in out.props)


# ---------------------------------------------------------------------------
# safety regressions (test/test_regressions.py)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("width,height", [(2, 2), (2, 64), (64, 2)])
def test_degenerate_axes(device, width, height):
    clip = blank(F420, width, height, length=2, color=(0.4, 0.1, -0.1))
    assert_finite(core.lgcr.Recon(clip).get_frame(0), f"Recon {width}x{height}")
    assert_finite(core.lgcr.TRecon(clip).get_frame(0), f"TRecon {width}x{height}")


def test_invalid_arguments_are_recoverable(device):
    clip = blank(F420, 16, 16)
    for kernel in ("lanczos", "jinc"):
        for taps in (-1, 0, 65):
            with pytest.raises(Error, match=r"1\.\.64"):
                core.lgcr.Recon(clip, kernel=kernel, taps=taps).get_frame(0)
    assert_finite(core.lgcr.Recon(clip, kernel="lanczos", taps=3).get_frame(0), "after invalid taps")
    for algo in (1, 3, 5):
        with pytest.raises(Error, match=r"2 \(guided\), 4 \(selector\) or 6"):
            core.lgcr.Recon(clip, algo=algo).get_frame(0)
    assert_finite(core.lgcr.Recon(clip, algo=2).get_frame(0), "after retired algorithms")


def test_temporal_chroma_gate(device):
    width = height = 64
    yy, xx = np.mgrid[0:height, 0:width]
    y = (0.35 + 0.08 * np.sin(xx * 0.31) * np.cos(yy * 0.27)).astype(np.float32)

    def fill(n, planes):
        planes[0][:] = y
        planes[1][:] = 0.0 if n == 0 else 0.2
        planes[2][:] = 0.0 if n == 0 else -0.2

    clip = from_fill(F420, width, height, 2, fill)
    single = core.lgcr.Recon(clip, strength=0.8, algo=2, sparse=0).get_frame(0)
    temporal = core.lgcr.TRecon(clip, strength=0.8, trad=3).get_frame(0)
    for p in range(3):
        assert np.array_equal(single[p], temporal[p]), f"chroma-inconsistent neighbour affected plane {p}"


def test_boundary_deduplication(device):
    rng = np.random.default_rng(19)
    width = height = 64
    y = rng.random((height, width), dtype=np.float32)
    u = rng.uniform(-0.3, 0.3, (height // 2, width // 2)).astype(np.float32)
    v = rng.uniform(-0.3, 0.3, (height // 2, width // 2)).astype(np.float32)

    def fill(n, planes):
        planes[0][:], planes[1][:], planes[2][:] = y, u, v

    clip = from_fill(F420, width, height, 2, fill)
    for n in (0, 1):
        reference = arrays(core.lgcr.TRecon(clip, trad=1).get_frame(n))
        for trad in (2, 3):
            candidate = core.lgcr.TRecon(clip, trad=trad).get_frame(n)
            for p in range(3):
                assert np.array_equal(reference[p], candidate[p]), f"frame {n} changed with trad={trad}, plane {p}"


def test_trecon_444_identity(device):
    rng = np.random.default_rng(23)
    width, height = 48, 32
    expected = [rng.uniform(-0.4 if p else 0.0, 0.4 if p else 1.0, (height, width)).astype(np.float32)
                for p in range(3)]
    clip = clip_from_planes([tuple(expected)], F444)
    output = core.lgcr.TRecon(clip, strength=0).get_frame(0)
    for p in range(3):
        assert np.array_equal(expected[p], output[p]), f"444 identity plane {p}"


# ---------------------------------------------------------------------------
# collaborative chroma switch (test/test_bm.py)
# ---------------------------------------------------------------------------

def make_noisy_clip(width=64, height=64, length=3):
    rng = np.random.default_rng(107)
    noise_u = rng.normal(0.0, 0.006, (height, width)).astype(np.float32)
    noise_v = rng.normal(0.0, 0.006, (height, width)).astype(np.float32)

    def fill(n, planes):
        planes[0][:] = 0.45
        planes[1][:] = 0.08 + noise_u
        planes[2][:] = -0.06 + noise_v

    return from_fill(F444, width, height, length, fill)


def chroma_variance(planes):
    return float(np.mean([np.var(np.asarray(planes[p]).astype(np.float64)) for p in (1, 2)]))


def test_bm_default_off_and_denoising(device):
    clip = make_noisy_clip()
    implicit = arrays(core.lgcr.Recon(clip, strength=0).get_frame(1))
    disabled = arrays(core.lgcr.Recon(clip, strength=0, bm=False).get_frame(1))
    enabled = core.lgcr.Recon(clip, strength=0, bm=True).get_frame(1)
    for p in range(3):
        assert np.array_equal(implicit[p], disabled[p])
        assert np.isfinite(enabled[p]).all()
    assert np.array_equal(implicit[0], enabled[0]), "collaborative stage modified luma"
    before, after = chroma_variance(disabled), chroma_variance(enabled)
    print(f"Recon bm noise variance: {before:.8f} -> {after:.8f}")
    assert after < 0.85 * before


def test_bm_constant_preservation(device):
    clip = blank(F444, 40, 32, color=(0.4, 0.1, -0.1))
    disabled = arrays(core.lgcr.Recon(clip, strength=0, bm=False).get_frame(0))
    enabled = arrays(core.lgcr.Recon(clip, strength=0, bm=True).get_frame(0))
    error = max(float(np.max(np.abs(disabled[p] - enabled[p]))) for p in (1, 2))
    assert error <= 1e-6


def test_bm_trecon_switch(device):
    clip = make_noisy_clip()
    implicit = arrays(core.lgcr.TRecon(clip, strength=0, trad=1).get_frame(1))
    disabled = arrays(core.lgcr.TRecon(clip, strength=0, trad=1, bm=False).get_frame(1))
    enabled = core.lgcr.TRecon(clip, strength=0, trad=1, bm=True).get_frame(1)
    for p in range(3):
        assert np.array_equal(implicit[p], disabled[p])
        assert np.isfinite(enabled[p]).all()
    assert chroma_variance(enabled) < 0.85 * chroma_variance(disabled)


def test_bm_edge_preservation(device):
    # The reference script derives the 4:2:0 source with zimg's bilinear
    # resize (not available here); a 2x2 box average of the same step edge
    # keeps the property under test: bm does not blur a clean co-edge.
    width = height = 64
    yy, xx = np.mgrid[0:height, 0:width]
    side = xx >= width // 2
    expected = (np.where(side, 0.58, 0.24).astype(np.float32), np.where(side, 0.22, -0.18).astype(np.float32),
                np.where(side, -0.20, 0.16).astype(np.float32))

    def box(a):
        return a.reshape(height // 2, 2, width // 2, 2).mean(axis=(1, 3)).astype(np.float32)

    source = clip_from_planes([(expected[0].copy(), box(expected[1]), box(expected[2]))], F420)
    disabled = core.lgcr.Recon(source, algo=2, sparse=0, bm=False).get_frame(0)
    enabled = core.lgcr.Recon(source, algo=2, sparse=0, bm=True).get_frame(0)
    band = np.abs(xx - width // 2) <= 4

    def edge_mae(frame):
        return float(np.mean([np.abs(frame[p].astype(np.float64) - expected[p])[band].mean() for p in (1, 2)]))

    before, after = edge_mae(disabled), edge_mae(enabled)
    print(f"bm co-edge MAE: {before:.8f} -> {after:.8f}")
    assert after <= before + 1e-5


def test_bm_scaled_and_integer_paths(device):
    clip = make_noisy_clip(width=48, height=32, length=1)
    scaled = core.lgcr.Recon(clip, width=80, height=56, bm=True).get_frame(0)
    assert scaled.width == 80 and scaled.height == 56
    assert_finite(scaled, "scaled bm")
    f420_8 = query_video_format(YUV, INTEGER, 8, 1, 1)
    integer = blank(f420_8, 48, 32, length=3, color=(64, 128, 128))
    for name, node in (("Recon", core.lgcr.Recon(integer, bm=True)), ("TRecon", core.lgcr.TRecon(integer, bm=True))):
        frame = node.get_frame(1)
        values = [int(frame[p][16, 24]) for p in range(3)]
        assert all(abs(a - b) <= 1 for a, b in zip(values, (64, 128, 128))), f"{name}: {values}"


# ---------------------------------------------------------------------------
# siting from frame properties, oracle parity through the plugin
# ---------------------------------------------------------------------------

def _random_420p10(width, height, length, seed):
    rng = np.random.default_rng(seed)
    fmt = query_video_format(YUV, INTEGER, 10, 1, 1)
    yy, xx = np.mgrid[0:height, 0:width]
    frames = []
    for n in range(length):
        y = (512 + 300 * np.sin(xx * 0.21 + n) * np.cos(yy * 0.17) + rng.normal(0, 8, (height, width)))
        u = 512 + rng.normal(0, 60, (height // 2, width // 2))
        v = 512 + rng.normal(0, 60, (height // 2, width // 2))
        frames.append(tuple(np.clip(np.rint(a), 0, 1023).astype(np.uint16) for a in (y, u, v)))
    return fmt, frames


@pytest.mark.parametrize("chroma_location,loc,expect", [
    (None, None, (-0.5, 0.0)),
    (2, None, (-0.5, -0.5)),     # top-left
    (1, None, (0.0, 0.0)),       # center
    (5, None, (0.0, 0.5)),       # bottom
    (2, "center", (0.0, 0.0)),   # an explicit loc wins over the property
])
def test_siting_from_props_matches_oracle(device, chroma_location, loc, expect):
    width, height = 96, 64
    fmt, frames = _random_420p10(width, height, 1, seed=41)
    props = {} if chroma_location is None else {"_ChromaLocation": chroma_location}
    clip = clip_from_planes(frames, fmt, props)
    kw = {} if loc is None else {"loc": loc}
    got = core.lgcr.Recon(clip, width=144, height=96, **kw).get_frame(0)
    ref = oracle.recon_frame(oracle.make_params(144, 96, shift_x=expect[0], shift_y=expect[1]), *frames[0], bits=10)
    for p in range(3):
        cnt, mx = compare_planes(ref[p], got[p], False)
        assert cnt == 0, f"plane {p}: mismatches={cnt} max_abs={mx:g}"


def test_trecon_and_sharpen_match_oracle(device):
    width, height = 80, 48
    fmt, frames = _random_420p10(width, height, 4, seed=43)
    clip = clip_from_planes(frames, fmt)
    n = 2
    got = core.lgcr.TRecon(clip, trad=1).get_frame(n)
    ids = plugin.temporal_frame_numbers(n, 3, 1)
    ref = oracle.trecon_frame(oracle.trecon_params(), frames[n], [frames[i] for i in ids], bits=10)
    for p in range(3):
        cnt, mx = compare_planes(ref[p], got[p], False)
        assert cnt == 0, f"TRecon plane {p}: mismatches={cnt} max_abs={mx:g}"
    sharp = core.lgcr.Sharpen(clip, alpha=0.5).get_frame(n)
    ref = oracle.sharpen_frame(oracle.sharpen_params(alpha=0.5), *frames[n], bits=10)
    for p in range(3):
        cnt, mx = compare_planes(ref[p], sharp[p], False)
        assert cnt == 0, f"Sharpen plane {p}: mismatches={cnt} max_abs={mx:g}"
    assert sharp.format == fmt


def _oracle_axis(src_n, dst_n, r, shift, params):
    import ctypes as C
    import oracle
    a = oracle.Axis()
    assert oracle.lib().or_build_axis(src_n, dst_n, r, shift, C.byref(params), C.byref(a)) == 0
    act = np.ctypeslib.as_array(a.active, shape=(dst_n,)).astype(np.int64).copy()
    start = np.ctypeslib.as_array(a.start, shape=(dst_n,)).copy()
    am = np.ctypeslib.as_array(a.am, shape=(dst_n * a.sup,)).reshape(dst_n, a.sup).copy()
    sup = a.sup
    oracle.lib().or_free_axis(C.byref(a))
    return act, start, am, sup


@pytest.mark.parametrize("sparse", [1, 0])
def test_profile_tap_counts(device, monkeypatch, sparse):
    """_LGCR_plainChroma_taps, _LGCR_reconstructChroma_taps and
    _LGCR_taps_visited count what the reference visits: the separable plain
    base computes each chroma row through a ring of `sup` rows and visits
    the active horizontal taps per computed row plus the active vertical
    taps per output pixel (recon.cpp:1453-1459); the guided pass visits
    activeX * activeY taps per workset pixel (recon.cpp:253-257), here taken
    from the oracle's workset and axis tables."""
    import oracle
    from lgcr_amd import synth
    monkeypatch.setenv("LGCR_PROFILE", "1")
    w, h = 128, 72
    y, u, v = synth.clip_frame(w, h, bits=10, seed=4, polygons=10)
    fmt = query_video_format(YUV, INTEGER, 10, 1, 1)
    src = clip_from_planes([(y, u, v)], fmt)
    f = core.lgcr.Recon(src, sparse=sparse).get_frame(0)
    params = oracle.make_params(w, h, sparse=sparse)
    oracle.recon_frame(params, y, u, v, bits=10)
    dbg = oracle.last_debug()
    ax, _, _, _ = _oracle_axis(w, w, 2.0, -0.5, params)
    ay, ystart, yam, ysup = _oracle_axis(h, h, 2.0, 0.0, params)
    ch = h // 2
    ring, cached, rows = min(ch, ysup), {}, 0
    for oy in range(h):
        for j in range(ysup):
            if yam[oy, j] == 0.0:
                continue
            s = min(max(int(ystart[oy]) + j, 0), ch - 1)
            if cached.get(s % ring) != s:
                cached[s % ring] = s
                rows += 1
    plain = rows * int(ax.sum()) + w * int(ay.sum())
    if sparse:
        idx = dbg["indices"].astype(np.int64)
        guided = int((ax[idx % w] * ay[idx // w]).sum())
        assert 0 < guided < int(ax.sum()) * int(ay.sum())
    else:
        guided = int(ax.sum()) * int(ay.sum())
    assert f.props["_LGCR_plainChroma_taps"] == plain
    assert f.props["_LGCR_reconstructChroma_taps"] == guided
    assert f.props["_LGCR_taps_visited"] == plain + guided
    assert f.props["_LGCR_cpu_guided_tap_accumulation_us"] == 0.0
