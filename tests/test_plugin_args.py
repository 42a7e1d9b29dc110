"""Argument handling of the plugin-surface mirror (lgcrCreate,
tReconCreate, sharpenCreate): defaults and the reference's error messages,
checked on the CPU before any device work (test/test_regressions.py:39-63
match the same substrings)."""
import numpy as np
import pytest

from lgcr_amd import abi, plugin
from lgcr_amd.plugin import Error, FLOAT, INTEGER, YUV, clip_from_planes, query_video_format


def clip(ss_w=1, ss_h=1, sample_type=INTEGER, bits=8, w=16, h=16, n=1):
    fmt = query_video_format(YUV, sample_type, bits, ss_w, ss_h)
    dt = fmt.dtype
    frames = [(np.zeros((h, w), dt), np.zeros((h >> ss_h, w >> ss_w), dt), np.zeros((h >> ss_h, w >> ss_w), dt))
              for _ in range(n)]
    return clip_from_planes(frames, fmt)


@pytest.mark.parametrize("kernel", ["lanczos", "jinc"])
@pytest.mark.parametrize("taps", [-1, 0, 65])
def test_invalid_taps(kernel, taps):
    with pytest.raises(Error, match=r"1\.\.64"):
        plugin.recon_params(clip(), kernel=kernel, taps=taps)


@pytest.mark.parametrize("algo", [1, 3, 5])
def test_retired_algorithms(algo):
    with pytest.raises(Error, match=r"2 \(guided\), 4 \(selector\) or 6"):
        plugin.recon_params(clip(), algo=algo)


@pytest.mark.parametrize("kw,msg", [
    ({"kernel": "gauss"}, 'kernel must be "bilinear"'),
    ({"loc": "top"}, 'loc must be "left" or "center"'),
    ({"width": 0}, "invalid width/height"),
    ({"strength": 1.5}, "invalid parameter range"),
    ({"bp": -0.1}, "invalid parameter range"),
    ({"reg": 0.0}, "invalid parameter range"),
    ({"qgate": 2.0}, "invalid parameter range"),
    ({"gamma": 1.0}, "argument gamma is not expected"),
])
def test_recon_errors(kw, msg):
    with pytest.raises(Error, match=msg):
        plugin.recon_params(clip(), **kw)


def test_recon_format_errors():
    with pytest.raises(Error, match="only 4:4:4 / 4:2:2 / 4:2:0"):
        plugin.recon_params(clip(ss_w=2, ss_h=2))
    with pytest.raises(Error, match="planar YUV"):
        plugin.recon_params(clip(sample_type=FLOAT, bits=16))


def test_recon_defaults_match_reference():
    p = plugin.recon_params(clip())
    assert (p.kernel, p.kp1, p.support, p.radial) == (abi.KERNELS["lanczos"], 3.0, 3.0, 0)
    assert (p.strength, p.sigma, p.sratio, p.sdb, p.stretch, p.gsigma, p.rescue) == (0.8, 0.01, 0.15, 3.0, 1.0,
                                                                                      2.5, 1.0)
    assert (p.algo, p.ridge, p.cedge, p.sparse, p.bm) == (2, 1, 0, 1, 0)
    assert (p.ms, p.qgate, p.ar_margin, p.reg, p.bp) == (1.0, 1.0, 0.0, 0.005, 0.0)
    assert (p.shift_x, p.shift_y, p.loc_set, p.out_w, p.out_h) == (-0.5, 0.0, 0, 16, 16)
    # siting is dropped on non-subsampled axes (plugin.cpp:1130-1133)
    assert plugin.recon_params(clip(ss_w=0, ss_h=0)).shift_x == 0.0
    q = plugin.recon_params(clip(), kernel="bicubic")
    assert (q.kp1, q.kp2, q.support) == (0.0, 0.6, 2.0)
    assert plugin.recon_params(clip(), loc="center").loc_set == 1


def test_trecon_errors_and_defaults():
    with pytest.raises(Error, match="invalid parameter range"):
        plugin.trecon_params(clip(), trad=9)
    with pytest.raises(Error, match="invalid parameter range"):
        plugin.trecon_params(clip(), tsad=0.0)
    with pytest.raises(Error, match="planar YUV 444/422/420"):
        plugin.trecon_params(clip(ss_w=2))
    with pytest.raises(Error, match="argument loc is not expected"):
        plugin.trecon_params(clip(), loc="left")
    p = plugin.trecon_params(clip())
    assert (p.trad, p.tsearch, p.tsad, p.sparse, p.algo, p.kp1) == (1, 6, 0.02, 0, 2, 3.0)


def test_sharpen_errors():
    with pytest.raises(Error, match="sigma/sratio/gspatial must be > 0"):
        plugin.Sharpen(clip(), sigma=0.0)
    with pytest.raises(Error, match="planar YUV only"):
        plugin.Sharpen(clip(sample_type=FLOAT, bits=16))


def test_siting_rules():
    fmt420 = query_video_format(YUV, INTEGER, 8, 1, 1)
    fmt444 = query_video_format(YUV, INTEGER, 8, 0, 0)
    s = plugin._siting
    assert s({"_ChromaLocation": 1}, fmt420, -0.5, 0.0, False) == (0.0, 0.0)
    assert s({"_ChromaLocation": 2}, fmt420, -0.5, 0.0, False) == (-0.5, -0.5)
    assert s({"_ChromaLocation": 5}, fmt420, -0.5, 0.0, False) == (0.0, 0.5)
    assert s({"_ChromaLocation": 1}, fmt420, -0.5, 0.0, True) == (-0.5, 0.0)
    assert s({"_ChromaLocation": 2}, fmt444, -0.5, 0.0, False) == (0.0, 0.0)
    assert s({}, fmt420, -0.5, 0.0, False) == (-0.5, 0.0)
    assert plugin.temporal_frame_numbers(0, 4, 2) == [1, 2]


def test_downsample_argument_errors():
    """downsampleCreate's checks and messages (downsample.cpp:816-893;
    test/test_downsample.py:341-362), raised before any device work."""
    f444 = plugin.query_video_format(plugin.YUV, plugin.FLOAT, 32, 0, 0)
    f420 = plugin.query_video_format(plugin.YUV, plugin.FLOAT, 32, 1, 1)
    rgb = plugin.query_video_format("RGB", plugin.FLOAT, 32, 0, 0)
    p444 = [np.zeros((16, 16), np.float32)] * 3
    odd = plugin.clip_from_planes([[np.zeros((14, 17), np.float32)] * 3], f444)
    sub = plugin.clip_from_planes([[np.zeros((16, 16), np.float32), np.zeros((8, 8), np.float32),
                                    np.zeros((8, 8), np.float32)]], f420)
    valid = plugin.clip_from_planes([p444], f444)
    cases = [
        (lambda: plugin.core.lgcr.Downsample(plugin.clip_from_planes([p444], rgb)), "input must be planar YUV444"),
        (lambda: plugin.core.lgcr.Downsample(sub), "input must be planar YUV444"),
        (lambda: plugin.core.lgcr.Downsample(odd), "constant even input width and height are required"),
        (lambda: plugin.core.lgcr.Downsample(valid, quality=-1), "quality must be 0, 1, or 2"),
        (lambda: plugin.core.lgcr.Downsample(valid, quality=3), "quality must be 0, 1, or 2"),
        (lambda: plugin.core.lgcr.Downsample(valid, kernel="lanczos"), "kernel must be"),
        (lambda: plugin.core.lgcr.Downsample(valid, strength=-0.01), "strength must be in the range 0..1"),
        (lambda: plugin.core.lgcr.Downsample(valid, strength=1.01), "strength must be in the range 0..1"),
        (lambda: plugin.core.lgcr.Downsample(valid, strength=float("nan")), "strength must be in the range 0..1"),
        (lambda: plugin.core.lgcr.Downsample(valid, loc="middle"), "loc must be left, center"),
        (lambda: plugin.core.lgcr.Downsample(valid, sigma=1.0), "argument sigma is not expected"),
    ]
    for call, msg in cases:
        with pytest.raises(plugin.Error, match=msg):
            call()
