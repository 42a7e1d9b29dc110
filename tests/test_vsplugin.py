"""The compiled VapourSynth 4 plugin (liblgcr_vs.so) in a minimal VS4 host,
without a device: registration (src/plugin.cpp:1160-1184,
src/downsample.cpp:913-918), the create functions' argument errors, and a
clean error when no HIP device exists."""
import os

import numpy as np
import pytest

import vs4host
from lgcr_amd import paths

# the reference's registration strings
SIGNATURES = {
    "Recon": ("clip:vnode;width:int:opt;height:int:opt;kernel:data:opt;taps:int:opt;algo:int:opt;"
              "b:float:opt;c:float:opt;strength:float:opt;sigma:float:opt;sratio:float:opt;"
              "sdb:float:opt;stretch:float:opt;gsigma:float:opt;rescue:float:opt;ridge:int:opt;cedge:int:opt;"
              "ar:float:opt;reg:float:opt;loc:data:opt;sparse:int:opt;bp:float:opt;ms:float:opt;qgate:float:opt;"
              "bm:int:opt", "clip:vnode"),
    "Sharpen": ("clip:vnode;alpha:float:opt;sigma:float:opt;sratio:float:opt;gspatial:float:opt;ar:float:opt",
                "clip:vnode"),
    "TRecon": ("clip:vnode;strength:float:opt;sigma:float:opt;sratio:float:opt;sdb:float:opt;"
               "gsigma:float:opt;stretch:float:opt;ar:float:opt;ridge:int:opt;sparse:int:opt;ms:float:opt;"
               "trad:int:opt;tsearch:int:opt;tsad:float:opt;bm:int:opt", "clip:vnode"),
    "Downsample": ("clip:vnode;quality:int:opt;kernel:data:opt;strength:float:opt;loc:data:opt", "clip:vnode"),
}


@pytest.fixture()
def host():
    h = vs4host.Host()
    h.load(paths.VS_PLUGIN)
    yield h
    h.close()


def source(host, ss=(1, 1), bits=8, w=32, h=16, n=3):
    dt = np.uint8 if bits == 8 else np.uint16
    cw, ch = w >> ss[0], h >> ss[1]
    frames = [(np.full((h, w), 100, dt), np.full((ch, cw), 90, dt), np.full((ch, cw), 160, dt)) for _ in range(n)]
    return host.source(frames, bits, *ss)


def test_registration(host):
    assert host.plugin() == ("dev.bsflab.lgcr", "lgcr",
                             "Luma-guided chroma reconstruction with direction-aware kernels")
    assert host.functions() == SIGNATURES


@pytest.mark.parametrize("fn,args,msg", [
    ("Recon", {"taps": 0}, "LGCR: lanczos taps must be in the range 1..64"),
    ("Recon", {"kernel": "jinc", "taps": 65}, "LGCR: jinc taps must be in the range 1..64"),
    ("Recon", {"algo": 3}, "LGCR: algo must be 2 (guided), 4 (selector) or 6 (detail transfer)"),
    ("Recon", {"kernel": "box"}, 'LGCR: kernel must be "bilinear", "bicubic", "lanczos", "spline16", '
                                 '"spline36" or "jinc"'),
    ("Recon", {"strength": 1.5}, "LGCR: invalid parameter range"),
    ("Recon", {"loc": "top"}, 'LGCR: loc must be "left" or "center"'),
    ("Recon", {"width": -4}, "LGCR: invalid width/height"),
    ("Sharpen", {"sigma": 0.0}, "LGCR Sharpen: sigma/sratio/gspatial must be > 0"),
    ("TRecon", {"trad": 9}, "TRecon: invalid parameter range"),
    ("TRecon", {"tsad": 0.0}, "TRecon: invalid parameter range"),
])
def test_create_errors(host, fn, args, msg):
    with pytest.raises(vs4host.HostError) as e:
        host.invoke(fn, source(host), **args)
    assert msg in str(e.value)


def test_downsample_create_errors(host):
    with pytest.raises(vs4host.HostError, match="Downsample: input must be planar YUV444"):
        host.invoke("Downsample", source(host))
    s444 = source(host, ss=(0, 0))
    for args, msg in (({"quality": 3}, "Downsample: quality must be 0, 1, or 2"),
                      ({"kernel": "bicubic"}, "Downsample: kernel must be"),
                      ({"strength": 2.0}, "Downsample: strength must be in the range 0..1"),
                      ({"loc": "middle"}, "Downsample: loc must be left, center")):
        with pytest.raises(vs4host.HostError, match=msg):
            host.invoke("Downsample", s444, **args)
    with pytest.raises(vs4host.HostError, match="constant even input width and height"):
        host.invoke("Downsample", source(host, ss=(0, 0), w=31))


def test_no_device_is_a_clean_error(host):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    for fn, src in (("Recon", source(host)), ("TRecon", source(host)), ("Sharpen", source(host)),
                    ("Downsample", source(host, ss=(0, 0)))):
        with pytest.raises(vs4host.HostError, match="no HIP device"):
            host.invoke(fn, src)
