"""Frames through the compiled VapourSynth 4 plugin (liblgcr_vs.so) in the
minimal VS4 host, against the oracle: Recon (luma plane shared, siting from
_ChromaLocation, the prop removed), TRecon (neighbour requests), Sharpen,
Downsample (luma shared, _ChromaLocation set) and the profile properties."""
import numpy as np
import pytest

import oracle
import vs4host
from lgcr_amd import abi, paths, synth

pytestmark = pytest.mark.gpu


@pytest.fixture()
def host(device):
    abi.lib()   # torch's HIP runtime first, then the plugin binds to it
    h = vs4host.Host()
    h.load(paths.VS_PLUGIN)
    yield h
    h.close()


def frames(n, w, h, bits, ss=(1, 1), **kw):
    return [synth.clip_frame(w, h, ss_w=ss[0], ss_h=ss[1], bits=bits, seed=31, frame=k, motion=True, **kw)
            for k in range(n)]


def test_recon_through_plugin(host):
    src = frames(2, 160, 96, 10)
    node = host.source(src, 10, 1, 1, chroma_location=0)
    out = host.invoke("Recon", node, bp=0.5, strength=0.7)
    assert host.info(out)[:5] == [3, 0, 10, 0, 0]
    for n in range(2):
        got = host.frame(out, n)
        # user floats are narrowed like mapGetFloatSaturated
        ref = oracle.recon_frame(oracle.make_params(160, 96, bp=0.5, strength=float(np.float32(0.7))), *src[n],
                                 bits=10)
        for i in range(3):
            assert np.array_equal(got[i], ref[i]), (n, i)
        assert host.shares_plane(0, node, n)
        assert "_ChromaLocation" not in host.props()


def test_recon_siting_from_frame_props(host):
    # _ChromaLocation 1 (center) moves the horizontal siting to 0
    src = frames(1, 96, 64, 8)
    node = host.source(src, 8, 1, 1, chroma_location=1)
    got = host.frame(host.invoke("Recon", node), 0)
    ref = oracle.recon_frame(oracle.make_params(96, 64, shift_x=0.0), *src[0])
    assert all(np.array_equal(got[i], ref[i]) for i in range(3))


def test_trecon_through_plugin(host):
    src = frames(3, 128, 72, 10)
    node = host.source(src, 10, 1, 1)
    out = host.invoke("TRecon", node)
    got = host.frame(out, 1)
    ref = oracle.trecon_frame(oracle.trecon_params(), src[1], [src[0], src[2]], bits=10)
    assert all(np.array_equal(got[i], ref[i]) for i in range(3))


def test_sharpen_through_plugin(host):
    src = frames(1, 96, 64, 8)
    node = host.source(src, 8, 1, 1, chroma_location=0)
    got = host.frame(host.invoke("Sharpen", node, alpha=0.5), 0)
    ref = oracle.sharpen_frame(oracle.sharpen_params(alpha=0.5), *src[0])
    assert all(np.array_equal(got[i], ref[i]) for i in range(3))
    assert host.props()["_ChromaLocation"] == 0   # Sharpen keeps the prop


def test_downsample_through_plugin(host):
    src = frames(1, 96, 64, 10, ss=(0, 0))
    node = host.source(src, 10, 0, 0)
    out = host.invoke("Downsample", node, quality=2, loc="topleft")
    assert host.info(out)[:5] == [3, 0, 10, 1, 1]
    got = host.frame(out, 0)
    ref = oracle.downsample_frame(oracle.downsample_params(quality=2, loc="topleft"), *src[0], bits=10)
    assert all(np.array_equal(got[i], ref[i]) for i in range(3))
    assert host.shares_plane(0, node, 0)
    assert host.props()["_ChromaLocation"] == 2


def test_profile_props_through_plugin(host, monkeypatch):
    monkeypatch.setenv("LGCR_PROFILE", "1")
    src = frames(1, 128, 72, 10)
    node = host.source(src, 10, 1, 1)
    host.frame(host.invoke("Recon", node), 0)
    props = host.props()
    assert props["_LGCR_input_conversion_pixels"] == 128 * 72 + 2 * 64 * 36
    assert props["_LGCR_output_conversion_pixels"] == 2 * 128 * 72
    assert props["_LGCR_total_us"] > 0.0
    assert props["_LGCR_taps_visited"] == props["_LGCR_plainChroma_taps"] + props["_LGCR_reconstructChroma_taps"]
    assert props["_LGCR_reconstructChroma_taps"] > 0
    assert props["_LGCR_cpu_guided_metadata_us"] == 0.0
    assert 0.0 < props["_LGCR_sparse_active_ratio"] <= 1.0


def test_injected_launch_fault_is_a_frame_error(host, monkeypatch):
    """A failed launch is that frame's filter error (setFilterError); the
    next request of the same frame is served."""
    src = frames(2, 96, 64, 10)
    node = host.source(src, 10, 1, 1)
    out = host.invoke("Recon", node)
    host.frame(out, 0)
    monkeypatch.setenv("LGCR_INJECT_FAULT", "launch:1")
    with pytest.raises(vs4host.HostError, match=r"LGCR: kernel launch failed: injected"):
        host.frame(out, 1)
    got = host.frame(out, 1)
    ref = oracle.recon_frame(oracle.make_params(96, 64), *src[1], bits=10)
    assert all(np.array_equal(got[i], ref[i]) for i in range(3))
