"""GPU parity for TRecon (motion-compensated temporal taps) against the
oracle, through the C ABI."""
import ctypes as C

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


# (w, h, tsearch): the window kernel holds the block and its search window in
# LDS, four blocks per workgroup up to tsearch 23 (at 16x16 blocks), one up to
# 54; tsearch 60 takes the global-memory kernel.  98x62 has partial blocks on
# the right and bottom; 160x96 has 60 blocks, so the last four-block group is
# full, 98x62 has 28, 128x80 40 and 144x80 45 (a partial last group).
ME_CASES = [(160, 96, 0), (160, 96, 2), (160, 96, 6), (98, 62, 6), (144, 80, 6), (160, 96, 20), (160, 96, 30),
            (128, 80, 60)]


@pytest.mark.parametrize("w,h,search", ME_CASES, ids=[f"{w}x{h}-s{s}" for w, h, s in ME_CASES])
def test_trecon_motion_vectors(device, w, h, search):
    """Block vectors and confidences of the first neighbour, bit-exact against
    the oracle's blockMatch + applyChromaConsistency (algos.cpp:683-807).
    Float frames, so the converted planes are the input planes."""
    frames = [synth.clip_frame(w, h, bits=32, seed=5, frame=k, motion=True, polygons=10) for k in range(3)]
    tsad = 0.02
    clip = abi.Clip(device, abi.trecon_params(tsearch=search, tsad=tsad), abi.make_format(np.float32, w, h),
                    temporal=True)
    try:
        clip.trecon_host(frames[1], [frames[0], frames[2]])
        mv = clip.fetch("mv", np.int16)
        conf = clip.fetch("conf")
    finally:
        clip.close()
    L = oracle.lib()
    bw, bh = (w + 15) // 16, (h + 15) // 16
    nb = bw * bh
    rmv = np.zeros(2 * nb, np.int16)
    rconf = np.zeros(nb, np.float32)
    i16 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int16))
    cy, cu, cv = (np.ascontiguousarray(p, np.float32) for p in frames[1])
    ny, nu, nv = (np.ascontiguousarray(p, np.float32) for p in frames[0])
    obw, obh = C.c_int(), C.c_int()
    L.or_block_match(oracle.fptr(cy), oracle.fptr(ny), w, h, 16, search, tsad, i16(rmv), i16(rmv[nb:]),
                     oracle.fptr(rconf), C.byref(obw), C.byref(obh))
    L.or_chroma_consistency(oracle.fptr(cu), oracle.fptr(cv), oracle.fptr(nu), oracle.fptr(nv), w // 2, h // 2, w, h,
                            16, i16(rmv), i16(rmv[nb:]), bw, bh, oracle.fptr(rconf))
    moving = int(np.count_nonzero(rmv[:nb]) + np.count_nonzero(rmv[nb:]))
    print(f"{w}x{h} s{search}: blocks={nb} nonzero components={moving} "
          f"mv mismatches={int(np.count_nonzero(mv != rmv))} conf mismatches={int(np.count_nonzero(conf != rconf))}")
    assert np.array_equal(mv, rmv)
    assert np.array_equal(conf.view(np.uint32), rconf.view(np.uint32))
