"""Sharpen (sharpenPlane, algos.cpp:595-659; sharpenGetFrame,
plugin.cpp:14-90): oracle checks.

The reference ships no Sharpen test, so the oracle is pinned by what the
mathematics fixes: constants and alpha=0 are identities, a linear ramp has a
zero gated Laplacian, an isolated impulse gains exactly (1+alpha) before the
hull clamp, mirror symmetry, and with a constant guide the gate is 1 so the
filter is a fixed 5x5 kernel 1/(1+(i^2+j^2)/gspatial^2), checked against
scipy's convolution with edge replication.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle


def params(**kw):
    return oracle.sharpen_params(**kw)


def test_constant_plane_is_fixed_point():
    p = np.full((20, 30), 0.37, np.float32)
    assert np.array_equal(oracle.sharpen_plane(params(), p), p)


def test_alpha_zero_is_identity():
    rng = np.random.default_rng(1)
    p = rng.random((24, 33)).astype(np.float32)
    assert np.array_equal(oracle.sharpen_plane(params(alpha=0.0), p), p)


def test_linear_ramp_has_zero_laplacian():
    yy, xx = np.mgrid[0:32, 0:40]
    p = (0.01 * xx + 0.004 * yy).astype(np.float32)
    out = oracle.sharpen_plane(params(alpha=0.8, ar=-1.0), p)
    inner = (slice(2, -2), slice(2, -2))
    assert np.abs(out[inner] - p[inner]).max() < 2e-6


@pytest.mark.parametrize("alpha", [0.3, 0.7])
def test_impulse_gain(alpha):
    p = np.zeros((15, 15), np.float32)
    p[7, 7] = 0.5
    free = oracle.sharpen_plane(params(alpha=alpha, ar=-1.0), p)
    assert free[7, 7] == pytest.approx(0.5 * (1.0 + alpha), abs=1e-6)
    clamped = oracle.sharpen_plane(params(alpha=alpha, ar=0.0), p)
    assert clamped[7, 7] == np.float32(0.5)
    # the hull clamp keeps every output inside its 5x5 source range
    assert clamped.min() >= 0.0 and clamped.max() <= 0.5


def test_mirror_symmetry():
    rng = np.random.default_rng(2)
    p = rng.random((21, 27)).astype(np.float32)
    a = oracle.sharpen_plane(params(alpha=0.5), p)
    b = oracle.sharpen_plane(params(alpha=0.5), p[:, ::-1].copy())
    c = oracle.sharpen_plane(params(alpha=0.5), p.T.copy())
    assert np.abs(a[:, ::-1] - b).max() < 1e-6
    assert np.abs(a.T - c).max() < 1e-6


def test_constant_guide_is_fixed_kernel():
    rng = np.random.default_rng(3)
    c = rng.random((18, 26)).astype(np.float32) - 0.5
    lc = np.full_like(c, 0.3)
    gy = np.full((36, 52), 0.3, np.float32)
    g = 1.3
    out = oracle.sharpen_plane(params(alpha=0.4, gspatial=g, ar=-1.0), c, lc=lc, gy=gy, rw=2.0, rh=2.0, sx=-0.5)
    jj, ii = np.mgrid[-2:3, -2:3]
    k = 1.0 / (1.0 + (ii ** 2 + jj ** 2) / g ** 2)
    k[2, 2] = 0.0
    c64 = c.astype(np.float64)
    nb = ndimage.correlate(c64, k, mode="nearest")
    ref = c64 + 0.4 * (k.sum() * c64 - nb) / k.sum()
    assert np.abs(out - ref).max() < 2e-6


def test_frame_float_444_and_420():
    rng = np.random.default_rng(4)
    y = rng.random((32, 48)).astype(np.float32)
    u = (rng.random((32, 48)) - 0.5).astype(np.float32)
    v = (rng.random((32, 48)) - 0.5).astype(np.float32)
    p = params()
    o = oracle.sharpen_frame(p, y, u, v, ss_w=0, ss_h=0)
    # 4:4:4: every plane is self-guided
    for a, b in zip(o, (y, u, v)):
        assert np.array_equal(a, oracle.sharpen_plane(p, b))
    u2, v2 = u[::2, ::2].copy(), v[::2, ::2].copy()
    o2 = oracle.sharpen_frame(p, y, u2, v2, ss_w=1, ss_h=1)
    assert np.array_equal(o2[0], o[0])
    assert o2[1].shape == u2.shape and not np.array_equal(o2[1], u2)


def test_integer_constant_frame_unchanged():
    y = np.full((16, 24), 200, np.uint16)
    c = np.full((8, 12), 512, np.uint16)
    o = oracle.sharpen_frame(params(), y, c, c, bits=10)
    assert np.array_equal(o[0], y) and np.array_equal(o[1], c) and np.array_equal(o[2], c)
