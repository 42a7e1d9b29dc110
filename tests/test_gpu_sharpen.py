"""GPU parity for Sharpen against the oracle, through the C ABI."""
import numpy as np
import pytest

import oracle
from lgcr_amd import abi, synth
from parity import compare_planes

pytestmark = pytest.mark.gpu


def planes(w, h, ss_w, ss_h, bits, seed):
    """Synthetic scene; chroma subsampled by 2^ss with a box of that size."""
    y, u, v = synth.clip_frame(w, h, ss_w=0, ss_h=0, bits=bits, seed=seed, polygons=10)
    if ss_w or ss_h:
        fx, fy = 1 << ss_w, 1 << ss_h

        def sub(p):
            a = p.astype(np.float64).reshape(h // fy, fy, w // fx, fx).mean((1, 3))
            return a.astype(np.float32) if bits == 32 else np.round(a).astype(p.dtype)

        u, v = sub(u), sub(v)
    return y, u, v


CASES = [
    # (id, w, h, ss_w, ss_h, bits, params)
    ("p8-420", 96, 64, 1, 1, 8, {}),
    ("p10-420", 96, 64, 1, 1, 10, {}),
    ("p16-422", 96, 64, 1, 0, 16, {}),
    ("p8-444", 96, 64, 0, 0, 8, {}),
    ("ps-420", 96, 64, 1, 1, 32, {}),
    ("ps-444-noclamp", 96, 64, 0, 0, 32, {"ar": -1.0}),
    ("p8-440", 96, 64, 0, 1, 8, {}),
    ("p8-410", 96, 64, 2, 2, 8, {}),
    ("p8-420-strong", 96, 64, 1, 1, 8, {"alpha": 0.9, "sigma": 0.003, "sratio": 0.3, "gspatial": 2.0,
                                        "ar": 0.01}),
    ("p8-420-odd", 98, 46, 1, 1, 8, {}),
    ("p8-444-tiny", 3, 2, 0, 0, 8, {}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_sharpen_matrix(device, case):
    tag, w, h, ssw, ssh, bits, kw = case
    y, u, v = planes(w, h, ssw, ssh, bits, seed=11)
    b = None if bits in (8, 16, 32) else bits
    ref = oracle.sharpen_frame(oracle.sharpen_params(**kw), y, u, v, bits=b, ss_w=ssw, ss_h=ssh)
    clip = abi.SharpenClip(device, abi.sharpen_params(**kw), abi.make_format(y.dtype, w, h, ssw, ssh, b))
    try:
        got = clip.sharpen_host(y, u, v)
    finally:
        clip.close()
    worst = 0
    for i, name in enumerate("YUV"):
        cnt, mx = compare_planes(ref[i], got[i], bits == 32)
        print(f"sharpen {tag} {name}: mismatches={cnt} max_abs={mx:g}")
        worst = max(worst, cnt)
    assert worst == 0


def test_sharpen_1080p_p10(device):
    y, u, v = synth.clip_frame(1920, 1080, bits=10, seed=5)
    ref = oracle.sharpen_frame(oracle.sharpen_params(), y, u, v, bits=10)
    clip = abi.SharpenClip(device, abi.sharpen_params(), abi.make_format(np.uint16, 1920, 1080, bits=10))
    try:
        got = clip.sharpen_host(y, u, v)
    finally:
        clip.close()
    for i, name in enumerate("YUV"):
        cnt, mx = compare_planes(ref[i], got[i], False)
        print(f"sharpen 1080p {name}: mismatches={cnt} max_abs={mx:g}")
        assert cnt == 0


@pytest.mark.parametrize("kw,msg", [({"sigma": 0.0}, "sigma/sratio/gspatial must be > 0"),
                                    ({"gspatial": -1.0}, "sigma/sratio/gspatial must be > 0")])
def test_sharpen_create_errors(device, kw, msg):
    with pytest.raises(abi.LgcrError, match=msg):
        abi.SharpenClip(device, abi.sharpen_params(**kw), abi.make_format(np.uint8, 64, 32))
