"""GPU parity for TRecon (motion-compensated temporal taps) against the
oracle, through the C ABI."""
import numpy as np
import pytest

import oracle
from lgcr_amd import abi, synth
from parity import compare_planes

pytestmark = pytest.mark.gpu

CASES = [
    # (id, w, h, ss_w, ss_h, bits, frames, n, params)
    ("p8-420-trad1", 160, 96, 1, 1, 8, 3, 1, {}),
    ("p10-420-trad2", 160, 96, 1, 1, 10, 5, 2, {"trad": 2}),
    ("ps-420-trad1", 160, 96, 1, 1, 32, 3, 1, {}),
    ("p8-420-boundary", 160, 96, 1, 1, 8, 3, 0, {"trad": 3}),
    ("p8-422-trad1", 160, 96, 1, 0, 8, 3, 1, {}),
    ("p8-444-trad1", 160, 96, 0, 0, 8, 3, 1, {}),
    ("p8-420-sparse", 160, 96, 1, 1, 8, 3, 1, {"sparse": 1}),
    ("p8-420-tsearch2", 160, 96, 1, 1, 8, 3, 1, {"tsearch": 2, "tsad": 0.05}),
    ("p8-420-strength0", 160, 96, 1, 1, 8, 3, 1, {"strength": 0.0}),
    ("p8-420-noms-ar", 160, 96, 1, 1, 8, 3, 1, {"ms": 0.0, "ar": -1.0}),
    ("p8-420-odd", 98, 62, 1, 1, 8, 3, 1, {}),
    ("p8-420-bm", 160, 96, 1, 1, 8, 3, 1, {"bm": 1}),
    ("ps-444-bm-strength0", 64, 64, 0, 0, 32, 3, 1, {"bm": 1, "strength": 0.0}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_trecon_matrix(device, case):
    tag, w, h, ssw, ssh, bits, nf, n, params = case
    frames = [synth.clip_frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=9, frame=k, motion=True, polygons=10)
              for k in range(nf)]
    trad = params.get("trad", 1)
    nbr_ids = oracle.temporal_frame_numbers(n, nf - 1, trad)
    nbrs = [frames[k] for k in nbr_ids]
    okw = dict(params)
    if "ar" in okw:
        okw["ar"] = okw["ar"]
    ref = oracle.trecon_frame(oracle.trecon_params(**okw), frames[n], nbrs, bits=None if bits in (8, 16, 32) else bits,
                              ss_w=ssw, ss_h=ssh)
    gkw = dict(params)
    gp = abi.trecon_params(**gkw)
    gp.shift_x = -0.5 if ssw else 0.0
    clip = abi.Clip(device, gp, abi.make_format(frames[0][0].dtype, w, h, ssw, ssh, bits if bits == 10 else None),
                    temporal=True)
    try:
        got = clip.trecon_host(frames[n], nbrs)
    finally:
        clip.close()
    worst = 0
    for i, name in enumerate("YUV"):
        cnt, mx = compare_planes(ref[i], got[i], bits == 32)
        print(f"{tag} {name}: mismatches={cnt} max_abs={mx:g}")
        worst = max(worst, cnt)
    assert worst == 0


def test_trecon_1080p_p10(device):
    """Serving-config frame: 1920x1080 YUV420P10 with moving polygons, trad 1."""
    frames = [synth.clip_frame(1920, 1080, bits=10, seed=4, frame=k, motion=True) for k in range(3)]
    ref = oracle.trecon_frame(oracle.trecon_params(), frames[1], [frames[0], frames[2]], bits=10)
    clip = abi.Clip(device, abi.trecon_params(), abi.make_format(np.uint16, 1920, 1080, bits=10), temporal=True)
    try:
        got = clip.trecon_host(frames[1], [frames[0], frames[2]])
    finally:
        clip.close()
    for i, name in enumerate("YUV"):
        cnt, mx = compare_planes(ref[i], got[i], False)
        print(f"trecon-1080p {name}: mismatches={cnt} max_abs={mx:g}")
        assert cnt == 0
