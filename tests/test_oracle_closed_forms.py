"""Closed-form checks for the oracle parts that no reference artefact pins:
TRecon's block confidence, zero-motion snap and chroma-consistency factor
(src/algos.cpp:683-807), and the algo-2 chroma-edge (cedge) fade
(src/recon.cpp:379-407).  Each case is built so that the quantities the
formula consumes (SAD, block variance, chroma deltas, edge width) are known
exactly, so a wrong constant, exponent, median rule or sampling stride
fails."""
import ctypes as C

import numpy as np
import pytest

import oracle

L = oracle.lib()
TSAD = 0.02


def block_match(cur, nbr, search=6, block=16):
    h, w = cur.shape
    bw, bh = C.c_int(), C.c_int()
    nb = ((w + block - 1) // block) * ((h + block - 1) // block)
    mvx, mvy = np.zeros(nb, np.int16), np.zeros(nb, np.int16)
    conf = np.zeros(nb, np.float32)
    L.or_block_match(oracle.fptr(np.ascontiguousarray(cur, np.float32)),
                     oracle.fptr(np.ascontiguousarray(nbr, np.float32)), w, h, block, search, TSAD,
                     mvx.ctypes.data_as(C.POINTER(C.c_int16)), mvy.ctypes.data_as(C.POINTER(C.c_int16)),
                     oracle.fptr(conf), C.byref(bw), C.byref(bh))
    shape = (bh.value, bw.value)
    return mvx.reshape(shape), mvy.reshape(shape), conf.reshape(shape)


def sub_var(cur, by, bx, block=16):
    """Population variance of the block's samples on the stride-2 lattice."""
    b = cur[by * block:(by + 1) * block:2, bx * block:(bx + 1) * block:2].astype(np.float64)
    return float(np.mean(b * b) - np.mean(b) ** 2)


def test_block_confidence_uniqueness_term():
    # identical neighbour: the zero vector has SAD 0, so confidence is the
    # observability term alone, var / (var + tsad^2 / 4), with var over the
    # stride-2 samples of the block; texture std 0.01 puts it near 1/2
    rng = np.random.default_rng(1)
    cur = (0.5 + 0.01 * rng.standard_normal((64, 64))).astype(np.float32)
    mvx, mvy, conf = block_match(cur, cur)
    assert not mvx.any() and not mvy.any()
    for by in range(4):
        for bx in range(4):
            var = sub_var(cur, by, bx)
            want = var / (var + 0.25 * TSAD * TSAD)
            assert abs(conf[by, bx] - want) <= 2e-6 * want, (by, bx, conf[by, bx], want)
    assert 0.3 < conf.mean() < 0.7


@pytest.mark.parametrize("offset", [TSAD, 2 * TSAD, 0.5 * TSAD])
def test_block_confidence_match_term(offset):
    # neighbour = current + constant: the zero vector wins with SAD/npix =
    # offset, so confidence = 1 / (1 + (offset / tsad)^2) * uniqueness
    rng = np.random.default_rng(2)
    cur = rng.random((64, 64), dtype=np.float32)
    nbr = (cur + np.float32(offset)).astype(np.float32)
    mvx, mvy, conf = block_match(cur, nbr)
    assert not mvx.any() and not mvy.any()
    per = np.abs(nbr.astype(np.float64) - cur)  # the float differences the SAD sums
    for by in range(4):
        for bx in range(4):
            var = sub_var(cur, by, bx)
            sad_px = float(per[by * 16:(by + 1) * 16:2, bx * 16:(bx + 1) * 16:2].mean())
            want = 1.0 / (1.0 + (sad_px / TSAD) ** 2) * var / (var + 0.25 * TSAD * TSAD)
            assert abs(conf[by, bx] - want) <= 1e-5 * want, (by, bx, conf[by, bx], want)


@pytest.mark.parametrize("margin,snapped", [(0.09, True), (0.11, False)])
def test_zero_motion_snap_margin(margin, snapped):
    # the neighbour is the current frame moved 2 px right, so (2, 0) has SAD
    # 0; the texture is scaled so the zero vector's SAD exceeds it by
    # `margin` * tsad * npix: below 0.10 the vector snaps to zero
    rng = np.random.default_rng(3)
    t = rng.random((64, 96))
    blk = (slice(16, 32), slice(32, 48))
    sub = t[blk][::2, ::2]
    shifted = t[16:32:2, 30:46:2]
    s = float(np.abs(sub - shifted).sum())
    npix = sub.size
    cur = (t * (margin * TSAD * npix / s)).astype(np.float32)
    nbr = np.roll(cur, 2, axis=1)
    mvx, mvy, conf = block_match(cur, nbr)
    assert mvy[1, 2] == 0
    assert mvx[1, 2] == (0 if snapped else 2)


def consistency(cu, cv, nu, nv, block=16):
    ch, cw = cu.shape
    bw, bh = (2 * cw + block - 1) // block, (2 * ch + block - 1) // block
    conf = np.ones(bw * bh, np.float32)
    zeros = np.zeros(bw * bh, np.int16)
    p16 = C.POINTER(C.c_int16)
    L.or_chroma_consistency(*(oracle.fptr(np.ascontiguousarray(a, np.float32)) for a in (cu, cv, nu, nv)),
                            cw, ch, 2 * cw, 2 * ch, block, zeros.ctypes.data_as(p16), zeros.ctypes.data_as(p16),
                            bw, bh, oracle.fptr(conf))
    return conf.reshape(bh, bw)


@pytest.mark.parametrize("delta,factor", [(0.04, 0.5), (0.02, 16.0 / 17.0), (0.08, 1.0 / 17.0)])
def test_chroma_consistency_factor(delta, factor):
    # every chroma sample of the block differs by delta (in V, U differs by
    # less): the median is delta and the factor 1 / (1 + (delta / 0.04)^4)
    rng = np.random.default_rng(4)
    cu = rng.uniform(-0.3, 0.3, (16, 16)).astype(np.float32)
    cv = rng.uniform(-0.3, 0.3, (16, 16)).astype(np.float32)
    nu = (cu + np.float32(0.5 * delta)).astype(np.float32)
    nv = (cv - np.float32(delta)).astype(np.float32)
    conf = consistency(cu, cv, nu, nv)
    np.testing.assert_allclose(conf, factor, rtol=2e-5)


def test_chroma_consistency_even_median():
    # an 8x8-sample block (even count): half its deltas are 0.02 and half
    # 0.06, so the median is their mean, 0.04, and the factor 1/2
    rng = np.random.default_rng(5)
    cu = rng.uniform(-0.3, 0.3, (16, 16)).astype(np.float32)
    cv = cu.copy()
    d = np.where(np.arange(16)[None, :] % 2 == 0, 0.02, 0.06) * np.ones((16, 1))
    nu = (cu + d).astype(np.float32)
    conf = consistency(cu, cv, nu, cv)
    np.testing.assert_allclose(conf, 0.5, rtol=1e-4)


def cedge_scene(width):
    """64x48 YUV420PS: a vertical luma step at column 32, and chroma that
    steps from (-0.2, 0.25) to (0.2, -0.15) over `width` chroma samples
    centred on the same position (width 1 = a hard step)."""
    h, w = 48, 64
    y = np.where(np.arange(w)[None, :] >= 32, 0.7, 0.3) * np.ones((h, 1))
    cx = np.arange(w // 2) + 0.5 - 16.0            # chroma sample centres relative to the edge
    t = np.clip(cx / width + 0.5, 0.0, 1.0)
    u = (-0.2 + 0.4 * t)[None, :] * np.ones((h // 2, 1))
    v = (0.25 - 0.4 * t)[None, :] * np.ones((h // 2, 1))
    return [a.astype(np.float32) for a in (y, u, v)]


def recon(planes, **kw):
    # ms = 0: the mutual-structure fade would also drop guidance on soft
    # chroma under a hard luma step, which would hide the cedge fade
    return oracle.recon_frame(oracle.make_params(64, 48, sparse=0, ms=0.0, **kw), *planes)


def test_cedge_keeps_guidance_on_hard_chroma_steps():
    # width 1: the chroma range over the window equals its largest
    # along-normal step, W_C = 1 < 1.5, so the fade is 1 and cedge changes
    # nothing
    planes = cedge_scene(1)
    a, b = recon(planes, cedge=1), recon(planes, cedge=0)
    assert all(np.array_equal(a[i], b[i]) for i in range(3))
    assert np.abs(b[1] - recon(planes, strength=0.0)[1]).max() > 1e-2


def test_cedge_falls_back_to_plain_on_soft_chroma():
    # width 4: W_C = 4 > 2.2 wherever the window holds the whole ramp (the
    # columns next to the edge), the fade is 0 and the pixel is the plain
    # kernel's; without cedge the guided result differs there
    planes = cedge_scene(4)
    on, off, plain = recon(planes, cedge=1), recon(planes, cedge=0), recon(planes, strength=0.0)
    cols = slice(30, 34)
    for i in (1, 2):
        assert np.abs(on[i][:, cols] - plain[i][:, cols]).max() <= 1e-6
        assert np.abs(off[i][:, cols] - plain[i][:, cols]).max() > 1e-3


def test_cedge_partial_fade_at_width_two():
    # width 2: W_C = 2 lies inside the fade band (2.2 - 0.7, 2.2), so the
    # result at the edge is neither the unfaded guided one nor the plain one
    planes = cedge_scene(2)
    on, off, plain = recon(planes, cedge=1), recon(planes, cedge=0), recon(planes, strength=0.0)
    cols = slice(31, 33)
    assert np.abs(on[1][:, cols] - off[1][:, cols]).max() > 1e-4
    assert np.abs(on[1][:, cols] - plain[1][:, cols]).max() > 1e-4
