"""CPU pins for the oracle's Downsample (downsample.cpp): the reference's own
NumPy model (golden vectors written by tests/golden/make_downsample_vectors.py)
and the properties the reference's test_downsample.py checks."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "downsample_numpy.npz")
LOCS = tuple(oracle.DOWNSAMPLE_LOCATIONS)
KERNELS = tuple(oracle.DOWNSAMPLE_KERNELS)


def run(y, u, v, bits=None, **kw):
    return oracle.downsample_frame(oracle.downsample_params(**kw), y, u, v, bits=bits)


def _mirrored_tie(y, loc, cx, cy):
    """True when the secondary direction of the quality-2 classification at
    chroma sample (cx, cy) is an exact tie between the mirrored directions
    +-22.5 degrees: the 3-row luma tensor (downsample.cpp classify, central
    differences with clamped taps) has jxy == 0 exactly, so
    n.x^2 jxx + 2 n.x n.y jxy + n.y^2 jyy is equal for (c, s) and (c, -s)."""
    _, sx, sy = oracle.DOWNSAMPLE_LOCATIONS[loc]
    h, w = y.shape
    yd = y.astype(np.float64)

    def at(x, r):
        return yd[min(max(r, 0), h - 1), min(max(x, 0), w - 1)]

    ix = int(np.floor(2 * cx + 0.5 + sx + 0.5))
    iy = int(np.floor(2 * cy + 0.5 + sy + 0.5))
    jxx = jxy = jyy = 0.0
    for k in (-1, 0, 1):
        gx = 0.5 * (at(ix + 1, iy + k) - at(ix - 1, iy + k))
        gy = 0.5 * (at(ix, iy + k + 1) - at(ix, iy + k - 1))
        jxx, jxy, jyy = jxx + gx * gx, jxy + gx * gy, jyy + gy * gy
    return jxy == 0.0 and jxx > jyy


def test_numpy_reference_vectors():
    """Agreement with the reference's NumPy model.  Qualities 0 and 1: every
    sample within 2e-6; quality 2 on frame "a" (the frame of the reference's
    own test, test/test_downsample.py:141-170): every sample within 2e-5,
    that test's bound.  Frame "b" at quality 2: every sample within 2e-6
    except where the classification's secondary direction is an exact tie
    between the mirrored directions +-22.5 degrees (checked per sample from
    the luma tensor): the float32 comparisons of downsample.cpp (and the
    oracle) keep the first of two equal responses, while the float64
    direction table of the NumPy model breaks the tie the other way, so the
    secondary candidate differs there; those samples stay within 1e-4."""
    g = np.load(GOLDEN)
    n = 0
    worst = {0: 0.0, 1: 0.0, 2: 0.0}
    ties = 0
    for key in g.files:
        if not key.endswith("_u") or key.count("_") < 4:
            continue
        name, loc, kernel, q, s = key[:-2].split("_")
        quality, strength = int(q[1:]), float(s[1:])
        y, u, v = g[f"{name}_y"], g[f"{name}_u"], g[f"{name}_v"]
        _, ou, ov = run(y, u, v, quality=quality, kernel=kernel, strength=strength, loc=loc)
        err = np.maximum(np.abs(ou - g[key]), np.abs(ov - g[key[:-2] + "_v"]))
        worst[quality] = max(worst[quality], float(err.max()))
        if quality < 2:
            assert err.max() <= 2e-6, (key, float(err.max()))
        elif name == "a":
            assert err.max() <= 2e-5, (key, float(err.max()))
        else:
            for cy, cx in np.argwhere(err > 2e-6):
                assert _mirrored_tie(y, loc, cx, cy), (key, cy, cx, float(err[cy, cx]))
                assert err[cy, cx] <= 1e-4, (key, cy, cx, float(err[cy, cx]))
                ties += 1
        n += 1
    print(f"{n} vector pairs, worst |err| per quality {worst}, frame-b samples at a mirrored tie: {ties}")
    assert n == 2 * (len(LOCS) * len(KERNELS) * 3 + 3)
    assert ties <= 20


@pytest.mark.parametrize("bits", [8, 10, 16, 32])
def test_constant_planes(bits):
    """test/test_downsample.py:84-113: a constant frame stays constant for
    every kernel and quality; Y is untouched."""
    h, w = 26, 34
    if bits == 32:
        color = [0.375, 0.125, -0.20]
        planes = [np.full((h, w), c, np.float32) for c in color]
    else:
        m = (1 << bits) - 1
        color = [round(0.375 * m), round(0.625 * m), round(0.30 * m)]
        dt = np.uint8 if bits == 8 else np.uint16
        planes = [np.full((h, w), c, dt) for c in color]
    for kernel in KERNELS:
        for quality in (0, 1, 2):
            y, ou, ov = run(*planes, bits=None if bits in (8, 16, 32) else bits, kernel=kernel, quality=quality)
            assert np.array_equal(y, planes[0])
            for p, c in ((ou, color[1]), (ov, color[2])):
                err = float(np.max(np.abs(p.astype(np.float64) - c)))
                assert err <= (2e-7 if bits == 32 else 0), (bits, kernel, quality, err)


def edge_frame(h=24, w=32):
    yy, xx = np.mgrid[:h, :w]
    side = xx > yy + 3
    return ((0.2 + 0.5 * side).astype(np.float32),
            (-0.25 + 0.5 * side + 0.01 * np.sin(xx)).astype(np.float32),
            (0.20 - 0.4 * side + 0.01 * np.cos(yy)).astype(np.float32))


def test_strength_zero_and_continuity():
    """test/test_downsample.py:116-138: strength 0 is the same baseline for
    every quality, and strength 1e-5 stays within 1.1e-5 of it."""
    planes = edge_frame()
    for kernel in KERNELS:
        _, bu, bv = run(*planes, quality=0, kernel=kernel, strength=0.0)
        fu, fv = oracle.downsample_base(oracle.downsample_params(kernel=kernel), *planes)
        assert np.array_equal(bu, fu) and np.array_equal(bv, fv)
        for quality in (1, 2):
            _, cu, cv = run(*planes, quality=quality, kernel=kernel, strength=0.0)
            assert np.array_equal(bu, cu) and np.array_equal(bv, cv)
        _, nu, nv = run(*planes, quality=2, kernel=kernel, strength=1e-5)
        assert max(float(np.max(np.abs(nu - bu))), float(np.max(np.abs(nv - bv)))) <= 1.1e-5
        _, gu, gv = run(*planes, quality=2, kernel=kernel, strength=1.0)
        assert not np.array_equal(gu, bu)  # the edge does engage the guided path


def test_impulse_siting():
    """test/test_downsample.py:179-201: the binomial baseline of a chroma
    impulse has its centroid at the output siting of every location."""
    h = w = 32
    y = np.full((h, w), 0.4, np.float32)
    u = np.zeros((h, w), np.float32)
    v = np.zeros((h, w), np.float32)
    u[16, 16], v[16, 16] = 1.0, -1.0
    for loc, (_, sx, sy) in oracle.DOWNSAMPLE_LOCATIONS.items():
        _, ou, _ = run(y, u, v, loc=loc, kernel="binomial", strength=0.0)
        p = ou.astype(np.float64)
        yy, xx = np.mgrid[:p.shape[0], :p.shape[1]]
        assert abs((p * xx).sum() / p.sum() - (16 - 0.5 - sx) / 2.0) <= 1e-6
        assert abs((p * yy).sum() / p.sum() - (16 - 0.5 - sy) / 2.0) <= 1e-6


def test_binomial_weights_closed_form():
    """The binomial baseline of a separable impulse is the outer product of
    [1 4 6 4 1]/16 (integer phase) or [1 3 3 1]/8 (half phase) taps."""
    h = w = 24
    y = np.full((h, w), 0.4, np.float32)
    u = np.zeros((h, w), np.float32)
    u[12, 12] = 1.0
    _, ou, _ = run(y, u, u, loc="topleft", kernel="binomial", strength=0.0)  # shifts -0.5: 5 taps
    # output i samples source centres 2i; source 12 -> output 6 (weight 6/16), 5 and 7 (1/16)
    k5 = {5: 1 / 16, 6: 6 / 16, 7: 1 / 16}
    for r, wr in k5.items():
        for c, wc in k5.items():
            assert abs(ou[r, c] - wr * wc) <= 1e-7
    _, ou, _ = run(y, u, u, loc="center", kernel="binomial", strength=0.0)  # shifts 0: 4 taps
    # output i covers sources 2i-1..2i+2 with [1 3 3 1]/8: source 12 is the
    # last tap of output 5 (1/8) and the second of output 6 (3/8)
    k4 = {5: 1 / 8, 6: 3 / 8}
    for r, wr in k4.items():
        for c, wc in k4.items():
            assert abs(ou[r, c] - wr * wc) <= 1e-7


def test_hull_and_fallbacks():
    """test/test_downsample.py:208-270: the guided move never increases the
    baseline's overshoot of the local source hull, and isoluminant or
    luma-only content keeps the baseline exactly."""
    h, w = 32, 40
    yy, xx = np.mgrid[:h, :w]
    side = xx > yy + 4
    y = (0.2 + 0.55 * side).astype(np.float32)
    u = (-0.30 + 0.55 * side + 0.015 * np.sin(xx * 0.8)).astype(np.float32)
    v = (0.25 - 0.50 * side + 0.012 * np.cos(yy * 0.7)).astype(np.float32)
    for quality in (0, 1, 2):
        _, bu, bv = run(y, u, v, quality=quality, strength=0.0)
        _, gu, gv = run(y, u, v, quality=quality, strength=1.0)
        r = quality + 2
        for oy in range(h // 2):
            cy = 2 * oy + 0.5
            ys = np.clip(np.arange(int(np.ceil(cy - r)), int(np.floor(cy + r)) + 1), 0, h - 1)
            for ox in range(w // 2):
                cx = 2 * ox
                xs = np.clip(np.arange(int(np.ceil(cx - r)), int(np.floor(cx + r)) + 1), 0, w - 1)
                for base, guided, src in ((bu, gu, u), (bv, gv, v)):
                    win = src[np.ix_(ys, xs)]
                    lo, hi = float(win.min()), float(win.max())
                    before = max(lo - float(base[oy, ox]), float(base[oy, ox]) - hi, 0.0)
                    after = max(lo - float(guided[oy, ox]), float(guided[oy, ox]) - hi, 0.0)
                    assert after <= before + 2e-6
    iso = (np.full((h, w), 0.4, np.float32), (0.15 * np.sin(xx * 0.25)).astype(np.float32),
           (0.15 * np.cos(yy * 0.25)).astype(np.float32))
    luma_only = (y, np.full((h, w), 0.12, np.float32), np.full((h, w), -0.08, np.float32))
    for case in (iso, luma_only):
        for quality in (0, 1, 2):
            _, bu, bv = run(*case, quality=quality, strength=0.0)
            _, gu, gv = run(*case, quality=quality, strength=1.0)
            assert np.array_equal(bu, gu) and np.array_equal(bv, gv)


def test_boundary_sizes_finite():
    """test/test_downsample.py:341-347: tiny even frames give finite output."""
    rng = np.random.default_rng(3)
    for w, h in ((2, 2), (4, 2), (2, 6), (6, 4), (18, 14)):
        planes = [rng.random((h, w), dtype=np.float32) for _ in range(3)]
        for quality in (0, 1, 2):
            _, ou, ov = run(*planes, quality=quality)
            assert ou.shape == (h // 2, w // 2) and np.isfinite(ou).all() and np.isfinite(ov).all()


def test_integer_rounding():
    """Integer outputs are int(v * max + max / 2 + 0.5) clamped (writeRow,
    downsample.cpp:117-126): the baseline of a constant is the constant."""
    for bits, dt in ((8, np.uint8), (10, np.uint16)):
        m = (1 << bits) - 1
        planes = [np.full((8, 8), c, dt) for c in (0, m, 1)]
        _, ou, ov = run(*planes, bits=None if bits == 8 else bits, strength=0.0)
        assert (ou == m).all() and (ov == 1).all()
