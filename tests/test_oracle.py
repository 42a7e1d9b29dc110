"""Pins for the oracle (CPU restatement) before it is trusted as the checker.

The reference cannot be built here (its sources need VapourSynth4.h, which
the image does not have), so the restatement is pinned by
  * the reference's own known-answer tests (test/test_sparse_workset.cpp),
  * its invariants (test/test_backend_matrix.py, test/test_regressions.py),
  * closed forms and brute force on small inputs, computed independently
    here in numpy with the same float32 operation order where bit-exactness
    is expected.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from parity import frame

L = oracle.lib()


def kernel(name, x, p1=3.0, p2=0.6):
    return L.or_kernel_eval(oracle.KERNELS[name], x, p1, p2)


# --------------------------------------------------------------------------
# closed forms (lgcr.h:53-117)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("name,p1,p2", [("bilinear", 0, 0), ("bicubic", 0.0, 0.6), ("bicubic", 1 / 3, 1 / 3),
                                        ("spline16", 0, 0), ("spline36", 0, 0)])
def test_kernel_partition_of_unity(name, p1, p2):
    for t in np.linspace(0.0, 0.999, 37):
        s = sum(kernel(name, t + k, p1, p2) for k in range(-4, 5))
        assert abs(s - 1.0) < 1e-12, (name, t, s)


@pytest.mark.parametrize("name", ["bilinear", "spline16", "spline36", "lanczos"])
def test_kernel_interpolates(name):
    # interpolating kernels: 1 at 0, 0 at every other integer
    assert kernel(name, 0.0) == 1.0
    for k in range(1, 5):
        assert abs(kernel(name, float(k))) < 1e-15


def test_lanczos_closed_form():
    for x in (0.25, 0.5, 1.3, 2.7):
        a = 3.0
        expect = (math.sin(math.pi * x) / (math.pi * x)) * (math.sin(math.pi * x / a) / (math.pi * x / a))
        assert abs(kernel("lanczos", x, a) - expect) < 1e-15
    assert kernel("lanczos", 3.0, 3.0) == 0.0


def test_jinc_known_values():
    # J1(1), J1(2) and the first zero of J1 (3.8317059702075125)
    j = lambda x: L.or_jinc_windowed(x / math.pi, 1e9) * x / 2.0  # taps huge -> window ~1
    assert abs(j(1.0) - 0.44005058574493355) < 1e-12
    assert abs(j(2.0) - 0.5767248077568734) < 1e-12
    assert abs(L.or_jinc_windowed(3.8317059702075125 / math.pi, 3.0)) < 1e-12
    assert L.or_jinc_windowed(0.0, 3.0) == 1.0
    assert L.or_jinc_windowed(3.0, 3.0) == 0.0


def test_lut_lookup_linear():
    lut = np.array([0.0, 1.0, 4.0, 9.0], np.float32)
    p = oracle.fptr(lut)
    assert L.or_lut_lookup(p, 4, 1.5) == 2.5
    assert L.or_lut_lookup(p, 4, 10.0) == 9.0  # clamps to the last segment, fr <= 1
    assert L.or_lut_lookup(p, 4, 0.0) == 0.0  # distances are never negative


# --------------------------------------------------------------------------
# weight tables and axes (kernels.cpp:135-167, 337-404)
# --------------------------------------------------------------------------

def test_weights_normalised_and_identity():
    w = oracle.Weights()
    assert L.or_build_weights(1920, 3840, oracle.KERNELS["lanczos"], 4.0, 0.0, 4.0, 0.0, C.byref(w)) == 0
    arr = np.ctypeslib.as_array(w.w, shape=(w.n * w.sup,)).reshape(w.n, w.sup)
    assert w.sup == 8
    assert np.all(np.abs(arr.astype(np.float64).sum(1) - 1.0) < 1e-6)
    L.or_free_weights(C.byref(w))
    # same size: every output sits on a source sample -> a delta
    assert L.or_build_weights(64, 64, oracle.KERNELS["lanczos"], 3.0, 0.0, 3.0, 0.0, C.byref(w)) == 0
    arr = np.ctypeslib.as_array(w.w, shape=(w.n * w.sup,)).reshape(w.n, w.sup)
    start = np.ctypeslib.as_array(w.start, shape=(w.n,))
    for i in range(64):
        j = int(np.argmax(arr[i]))
        assert start[i] + j == i and abs(arr[i, j] - 1.0) < 1e-7
    L.or_free_weights(C.byref(w))


def test_axis_left_siting_phases():
    # 420 left siting, same size: pos = x/2 and lpos = x (kernels.cpp:358-360)
    p = oracle.make_params(64, 64)
    a = oracle.Axis()
    assert L.or_build_axis(64, 64, 2.0, -0.5, C.byref(p), C.byref(a)) == 0
    pos = np.ctypeslib.as_array(a.pos, shape=(64,))
    lpos = np.ctypeslib.as_array(a.lpos, shape=(64,))
    assert np.array_equal(pos, np.arange(64, dtype=np.float32) / 2)
    assert np.array_equal(lpos, np.arange(64, dtype=np.float32))
    act = np.ctypeslib.as_array(a.active, shape=(64,))
    # phase 0 hits 5 active taps (|d| < 3 excludes the edge), phase 1/2 six
    assert act[10] == 5 and act[11] == 6
    wsum = np.ctypeslib.as_array(a.wsum, shape=(64,))
    assert np.all(np.abs(wsum[4:-4] - 1.0) < 1e-6)
    L.or_free_axis(C.byref(a))


# --------------------------------------------------------------------------
# known-answer tests of test/test_sparse_workset.cpp
# --------------------------------------------------------------------------

def _axis_from_positions(outputs, source):
    """makeAxis of test_sparse_workset.cpp:11-24: chromaBilin only."""
    pos = (np.arange(outputs, dtype=np.float32) + np.float32(0.5)) * np.float32(source) / np.float32(outputs) \
        - np.float32(0.5)
    raw = np.floor(pos).astype(np.int32)
    i0 = np.zeros(outputs, np.int32) if source == 1 else np.clip(raw, 0, source - 2).astype(np.int32)
    a = oracle.Axis()
    a.n = outputs
    a.cb_i0 = i0.ctypes.data_as(C.POINTER(C.c_int))
    return a, i0


def test_kat_mask_and_spans():
    width, height, cw, ch = 8, 4, 4, 2
    mask = np.zeros((height, width), np.uint8)
    mask[1, 1:4] = 1
    mask[3, 5:8] = 1
    ax, kx = _axis_from_positions(width, cw)
    ay, ky = _axis_from_positions(height, ch)
    ws = oracle.Workset()
    L.or_build_workset(oracle.u8ptr(mask), width, height, width, height, C.byref(ax), C.byref(ay), cw, ch,
                       C.byref(ws))
    assert ws.n_indices == 6 and ws.n_spans == 2
    spans = np.ctypeslib.as_array(ws.spans, shape=(4,))
    assert list(spans) == [1, 4, 5, 8]
    so = np.ctypeslib.as_array(ws.span_row_offsets, shape=(height + 1,))
    io = np.ctypeslib.as_array(ws.index_row_offsets, shape=(height + 1,))
    assert so[0] == 0 and so[1] == 0 and so[2] == 1 and so[4] == 2
    assert io[0] == 0 and io[1] == 0 and io[2] == 3 and io[4] == 6
    idx = np.ctypeslib.as_array(ws.indices, shape=(6,))
    assert idx[0] == 9 and idx[-1] == 31
    assert not L.or_workset_output_dense(C.byref(ws))
    cm = np.ctypeslib.as_array(ws.chroma_mask, shape=(cw * ch,))
    assert ws.active_chroma == int(cm.sum())
    L.or_free_workset(C.byref(ws))
    full = np.ones_like(mask)
    L.or_build_workset(oracle.u8ptr(full), width, height, width, height, C.byref(ax), C.byref(ay), cw, ch,
                       C.byref(ws))
    assert L.or_workset_output_dense(C.byref(ws)) and L.or_workset_chroma_dense(C.byref(ws))
    L.or_free_workset(C.byref(ws))
    half = np.zeros_like(mask)
    half.reshape(-1)[: half.size // 2] = 1
    L.or_build_workset(oracle.u8ptr(half), width, height, width, height, C.byref(ax), C.byref(ay), cw, ch,
                       C.byref(ws))
    assert ws.n_indices * 2 == half.size and L.or_workset_output_dense(C.byref(ws))
    L.or_free_workset(C.byref(ws))


def test_kat_trust_dilation_impulse():
    # jxx impulse of 1 at (4,3), sigma 0.1 -> seed only there; radius 2 -> 5x5
    jxx = np.zeros((7, 9), np.float32)
    jxx[3, 4] = 1.0
    seed = (jxx + 0.0 > np.float32(0.1 * 0.1)).astype(np.uint8)
    out = np.zeros_like(seed)
    L.or_dilate(oracle.u8ptr(seed), 9, 7, 2, oracle.u8ptr(out))
    yy, xx = np.mgrid[0:7, 0:9]
    assert np.array_equal(out.astype(bool), (np.abs(xx - 4) <= 2) & (np.abs(yy - 3) <= 2))


def test_kat_fused_trust_seed():
    y = np.zeros((7, 9), np.float32)
    y[:, 4:] = 1.0
    jxx, jxy, jyy = (np.zeros_like(y) for _ in range(3))
    seed = np.zeros(y.shape, np.uint8)
    thr = np.float32(0.1 * 0.1)
    L.or_sobel_tensor(oracle.fptr(y), 9, 7, oracle.fptr(jxx), oracle.fptr(jxy), oracle.fptr(jyy),
                      oracle.u8ptr(seed), float(thr))
    fallback = (jxx + jyy > thr).astype(np.uint8)
    assert np.array_equal(seed, fallback)
    a, b = np.zeros_like(seed), np.zeros_like(seed)
    L.or_dilate(oracle.u8ptr(seed), 9, 7, 2, oracle.u8ptr(a))
    L.or_dilate(oracle.u8ptr(fallback), 9, 7, 2, oracle.u8ptr(b))
    assert np.array_equal(a, b)


def test_kat_mutual_roi():
    out_n, c_n = 16, 8
    ax, _ = _axis_from_positions(out_n, c_n)
    ay, _ = _axis_from_positions(out_n, c_n)
    mask = np.zeros((out_n, out_n), np.uint8)
    mask[3:13, 6:10] = 1
    ws = oracle.Workset()
    L.or_build_workset(oracle.u8ptr(mask), out_n, out_n, out_n, out_n, C.byref(ax), C.byref(ay), c_n, c_n,
                       C.byref(ws))
    cm = np.ctypeslib.as_array(ws.chroma_mask, shape=(c_n * c_n,)).copy()
    yy, xx = np.mgrid[0:c_n, 0:c_n]
    edge = (xx >= 4).astype(np.float32)
    lc = (np.float32(0.2) + np.float32(0.3) * edge + np.float32(0.01) * yy).astype(np.float32)
    u = (np.float32(-0.2) + np.float32(0.4) * edge).astype(np.float32)
    v = (np.float32(0.25) - np.float32(0.35) * edge).astype(np.float32)
    dense = np.zeros_like(lc)
    sparse = np.zeros_like(lc)
    L.or_mutual_gate(oracle.fptr(lc), oracle.fptr(u), oracle.fptr(v), c_n, c_n, None, 1, oracle.fptr(dense))
    L.or_mutual_gate(oracle.fptr(lc), oracle.fptr(u), oracle.fptr(v), c_n, c_n, oracle.u8ptr(cm),
                     int(L.or_workset_chroma_dense(C.byref(ws))), oracle.fptr(sparse))
    on = cm.reshape(c_n, c_n) != 0
    assert np.all(np.abs(dense[on] - sparse[on]) <= 1e-6)
    # co-located edges in luma and chroma: the gate confirms them
    assert dense[4, 3] > 0.5 and dense[4, 4] > 0.5
    L.or_free_workset(C.byref(ws))


# --------------------------------------------------------------------------
# brute force on small inputs (independent numpy, same float32 op order)
# --------------------------------------------------------------------------

def _np_sobel(p):
    h, w = p.shape
    P = np.pad(p, 1, mode="edge")
    tl, t, tr = P[:-2, :-2], P[:-2, 1:-1], P[:-2, 2:]
    l, r = P[1:-1, :-2], P[1:-1, 2:]
    bl, b, br = P[2:, :-2], P[2:, 1:-1], P[2:, 2:]
    two, e = np.float32(2), np.float32(0.125)
    gx = (((((tr + two * r) + br) - tl) - two * l) - bl) * e
    gy = (((((bl + two * b) + br) - tl) - two * t) - tr) * e
    return gx, gy


def test_brute_force_tensor():
    rng = np.random.default_rng(1)
    y = rng.random((13, 17), dtype=np.float32)
    gx, gy = _np_sobel(y)
    G = [np.pad(a, ((1, 1), (0, 0)), mode="edge") for a in (gx, gy)]
    vxx = (G[0][:-2] * G[0][:-2] + G[0][1:-1] * G[0][1:-1]) + G[0][2:] * G[0][2:]
    nine = np.float32(9.0)
    ninth = np.float32(1.0) / nine
    exp = np.empty_like(y)
    exp[:, 1:-1] = ninth * ((vxx[:, :-2] + vxx[:, 1:-1]) + vxx[:, 2:])
    exp[:, 0] = (np.float32(2) * vxx[:, 0] + vxx[:, 1]) / nine
    exp[:, -1] = (vxx[:, -2] + np.float32(2) * vxx[:, -1]) / nine
    jxx, jxy, jyy = (np.zeros_like(y) for _ in range(3))
    L.or_sobel_tensor(oracle.fptr(y), 17, 13, oracle.fptr(jxx), oracle.fptr(jxy), oracle.fptr(jyy), None, 0.0)
    assert np.array_equal(jxx, exp)


def test_brute_force_lc_map():
    rng = np.random.default_rng(2)
    y = rng.random((10, 12), dtype=np.float32)
    lc = np.zeros((5, 6), np.float32)
    L.or_lc_map(oracle.fptr(y), 12, 10, 6, 5, 2.0, 2.0, -0.5, 0.0, oracle.fptr(lc))
    for cy in range(5):
        for cx in range(6):
            cols = [c for c in (2 * cx - 1, 2 * cx) if c >= 0]
            s = 0.0
            for j in (2 * cy, 2 * cy + 1):
                for i in cols:
                    s += float(y[j, i])
            assert lc[cy, cx] == np.float32(s / (len(cols) * 2))


def test_brute_force_dilation():
    from scipy.ndimage import maximum_filter
    rng = np.random.default_rng(3)
    seed = (rng.random((40, 57)) > 0.985).astype(np.uint8)
    for r in (0, 1, 5, 14):
        out = np.zeros_like(seed)
        L.or_dilate(oracle.u8ptr(seed), 57, 40, r, oracle.u8ptr(out))
        exp = maximum_filter(seed, size=2 * r + 1, mode="constant", cval=0)
        assert np.array_equal(out, exp), r


def test_workset_indices_are_flatnonzero():
    y, u, v = frame(96, 64, bits=8, seed=4, polygons=6)
    oracle.recon_frame(oracle.make_params(96, 64), y, u, v)
    d = oracle.last_debug()
    assert np.array_equal(d["indices"], np.flatnonzero(d["mask"]).astype(np.uint32))
    # chroma mask = 2x2 bilinear footprints of active outputs (maps.cpp:385-401)
    m = d["mask"].reshape(64, 96)
    cm = np.zeros((32, 48), np.uint8)
    # left siting: pos_x = x / 2; centre rows: pos_y = (y + 0.5) / 2 - 0.5
    posx = np.float32(np.arange(96) / 2.0)
    posy = ((np.arange(64) + 0.5) / 2.0 - 0.5).astype(np.float32)
    ix = np.clip(np.floor(posx).astype(int), 0, 46)
    iy = np.clip(np.floor(posy).astype(int), 0, 30)
    for oy, ox in zip(*np.nonzero(m)):
        cy0, cx0 = iy[oy], ix[ox]
        for cy in (cy0, min(cy0 + 1, 31)):
            for cx in (cx0, min(cx0 + 1, 47)):
                cm[cy, cx] = 1
    assert np.array_equal(d["chroma_mask"].reshape(32, 48), cm)


# --------------------------------------------------------------------------
# invariants of the reference tests
# --------------------------------------------------------------------------

def _recon(y, u, v, ss_w=1, ss_h=1, **kw):
    h, w = y.shape
    p = oracle.make_params(w, h, **kw)
    out = oracle.recon_frame(p, y, u, v, ss_w=ss_w, ss_h=ss_h)
    return out, oracle.last_debug()


def test_444_strength0_identity():
    # test_regressions.py:119-136
    rng = np.random.default_rng(23)
    planes = [rng.uniform(-0.4 if p else 0.0, 0.4 if p else 1.0, (32, 48)).astype(np.float32) for p in range(3)]
    out, _ = _recon(*planes, ss_w=0, ss_h=0, strength=0.0)  # default left siting is dropped on 444
    for p in range(3):
        assert np.array_equal(out[p], planes[p])


def test_420_phase0_samples_are_exact():
    # top-left-sited 420, same size: output (2i, 2j) sits on chroma (i, j)
    y, u, v = frame(96, 64, bits=32, seed=5)
    out, _ = _recon(y, u, v, strength=0.0, shift_y=-0.5)
    assert np.max(np.abs(out[1][::2, ::2] - u)) <= 1e-7


@pytest.mark.parametrize("kw", [{}, {"strength": 0.0}, {"kernel": "bicubic"}, {"bp": 0.5}])
def test_constant_chroma_preserved(kw):
    y, _, _ = frame(64, 48, bits=32, seed=6)
    u = np.full((24, 32), np.float32(0.1))
    v = np.full((24, 32), np.float32(-0.1))
    out, _ = _recon(y, u, v, **kw)
    assert np.all(out[1] == np.float32(0.1)) and np.all(out[2] == np.float32(-0.1))


@pytest.mark.parametrize("kernel", ["lanczos", "bicubic", "spline16", "spline36", "bilinear"])
def test_sparse_equals_dense(kernel):
    # test_backend_matrix.py:104-111 (algo 2)
    y, u, v = frame(96, 64, bits=32, seed=8, polygons=10)
    dense, _ = _recon(y, u, v, kernel=kernel, sparse=0)
    sparse, _ = _recon(y, u, v, kernel=kernel, sparse=1)
    assert max(float(np.max(np.abs(dense[p] - sparse[p]))) for p in (1, 2)) <= 1e-6


@pytest.mark.parametrize("kernel", ["lanczos", "bicubic", "spline36"])
def test_strength_continuity(kernel):
    # test_backend_matrix.py:92-100
    y, u, v = frame(64, 48, bits=32, seed=9, polygons=8)
    zero, _ = _recon(y, u, v, kernel=kernel, strength=0.0, sparse=0)
    tiny, _ = _recon(y, u, v, kernel=kernel, strength=1e-6, sparse=0)
    full, _ = _recon(y, u, v, kernel=kernel, strength=1.0, sparse=0)
    step = max(float(np.max(np.abs(zero[p] - tiny[p]))) for p in (1, 2))
    total = max(float(np.max(np.abs(zero[p] - full[p]))) for p in (1, 2))
    assert total > 1e-3
    assert step <= total * 3e-6 + 2e-7


def test_back_projection_consistency():
    # bp = 1 returns the full residual: box2x2(out) == source chroma
    y, u, v = frame(96, 64, bits=32, seed=10)
    out, _ = _recon(y, u, v, bp=1.0)
    box = 0.25 * (out[1][0::2, 0::2] + out[1][1::2, 0::2] + out[1][0::2, 1::2] + out[1][1::2, 1::2])
    assert np.max(np.abs(box - u)) <= 2e-7


def _tri_420(p):
    """4:4:4 -> 4:2:0 the way a bilinear 2:1 downscale does it: [1/4, 1/2, 1/4]
    around even columns (left siting), [1/2, 1/2] over row pairs."""
    h, w = p.shape
    q = np.pad(p.astype(np.float64), ((0, 0), (1, 1)), mode="edge")
    x = 0.25 * q[:, 0:w:2] + 0.5 * q[:, 1:w + 1:2] + 0.25 * q[:, 2:w + 2:2]
    return (0.5 * (x[0::2] + x[1::2])).astype(np.float32)


def test_luma_untouched_and_guided_beats_plain():
    # test/test_lgcr.py:26-113: vertical edge (bottom half) and 45-degree edge
    # (top half) between two saturated colours; guided must beat the plain
    # kernel (ar=-1) in the bands next to both edges, and leave luma alone.
    H = W = 128
    c1, c2 = (0.30, -0.15, 0.40), (0.25, 0.40, -0.20)
    jj, ii = np.mgrid[0:H, 0:W]
    side = np.where(jj < H // 2, ii >= jj, ii >= W // 2)
    gt = [np.where(side, c2[p], c1[p]).astype(np.float32) for p in range(3)]
    y, u, v = gt[0], _tri_420(gt[1]), _tri_420(gt[2])
    plain, _ = _recon(y, u, v, strength=0.0, ar=-1.0)
    guided, _ = _recon(y, u, v)
    assert np.array_equal(guided[0], y)
    vert = (jj >= H // 2 + 4) & (jj < H - 4) & (np.abs(ii - W // 2) <= 3) & (ii != W // 2)
    diag = (jj >= 4) & (jj < H // 2 - 4) & (np.abs(ii - jj) > 0) & (np.abs(ii - jj) <= 2)
    for p in (1, 2):
        for band in (vert, diag):
            e_p = float(np.mean(np.abs(plain[p][band] - gt[p][band])))
            e_g = float(np.mean(np.abs(guided[p][band] - gt[p][band])))
            assert e_g < e_p, (p, e_g, e_p)


def test_integer_output_rounding():
    # floatToPlane (lgcr.h:1007-1010): int(v*scale + offset + 0.5f), clamped
    src = np.array([[-0.6, -0.5, -0.25, 0.0, 0.25, 0.4999, 0.5, 0.7]], np.float32)
    out = np.zeros(src.shape, np.uint16)
    L.or_float_to_u16.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_uint16), C.c_ssize_t, C.c_int, C.c_int,
                                  C.c_double, C.c_double, C.c_int, C.c_int]
    L.or_float_to_u16(oracle.fptr(src), out.ctypes.data_as(C.POINTER(C.c_uint16)), out.strides[0], 8, 1,
                      1023.0, 511.5, 0, 1023)
    f = np.float32
    exp = [max(0, min(1023, int(f(s) * f(1023.0) + f(511.5) + f(0.5)))) for s in src[0]]
    assert list(out[0]) == exp


# --------------------------------------------------------------------------
# algo 4 (LGF regression + selector)
# --------------------------------------------------------------------------

def test_lgf_recovers_affine_chroma():
    # chroma = a * Y + b exactly -> slope a, intercept b wherever the window
    # has luma variance well above eps (closed form of the regression)
    rng = np.random.default_rng(12)
    y = rng.random((32, 48), dtype=np.float32)
    ls = np.zeros((16, 24), np.float32)
    # point-sampled luma at 4:4:4 == Y itself: use rw = rh = 1 and no shift
    u = (np.float32(0.3) * y - np.float32(0.1)).astype(np.float32)
    v = (np.float32(-0.2) * y + np.float32(0.05)).astype(np.float32)
    out = np.zeros((6, 32, 48), np.float32)
    L.or_lgf_maps(oracle.fptr(y), 48, 32, 48, 32, 1.0, 1.0, 0.0, 0.0, oracle.fptr(u), oracle.fptr(v), 1e-12, 0, None,
                  oracle.fptr(out))
    assert np.max(np.abs(out[0] - 0.3)) < 1e-4 and np.max(np.abs(out[1] + 0.1)) < 1e-4
    assert np.max(np.abs(out[3] + 0.2)) < 1e-4 and np.max(np.abs(out[4] - 0.05)) < 1e-4
    assert np.all(out[2] > 0.999) and np.all(out[5] > 0.999)


def test_lgf_roi_only_changes_unselected_samples():
    rng = np.random.default_rng(13)
    y = rng.random((40, 64), dtype=np.float32)
    u = rng.random((20, 32), dtype=np.float32) - 0.5
    v = rng.random((20, 32), dtype=np.float32) - 0.5
    full = np.zeros((6, 20, 32), np.float32)
    part = np.zeros((6, 20, 32), np.float32)
    roi = np.zeros((20, 32), np.uint8)
    roi[5:9, 10:20] = 1
    args = (oracle.fptr(y), 64, 40, 32, 20, 2.0, 2.0, -0.5, 0.0, oracle.fptr(u), oracle.fptr(v), 2.5e-5, 1)
    L.or_lgf_maps(*args, None, oracle.fptr(full))
    L.or_lgf_maps(*args, oracle.u8ptr(roi), oracle.fptr(part))
    on = roi.astype(bool)
    assert np.array_equal(full[:, on], part[:, on])
    assert np.all(part[:, ~on] == 0)


@pytest.mark.parametrize("bits", [8, 32])
def test_algo4_sparse_equals_dense(bits):
    y, u, v = frame(96, 64, bits=bits, seed=14, polygons=10)
    dense, _ = _recon(y, u, v, algo=4, sparse=0)
    sparse, _ = _recon(y, u, v, algo=4, sparse=1)
    tol = 1e-6 if bits == 32 else 0
    assert max(float(np.max(np.abs(dense[p].astype(np.float64) - sparse[p].astype(np.float64)))) for p in (1, 2)) <= tol


def test_algo4_constant_chroma_preserved():
    y, _, _ = frame(64, 48, bits=32, seed=15)
    u = np.full((24, 32), np.float32(0.1))
    v = np.full((24, 32), np.float32(-0.1))
    out, _ = _recon(y, u, v, algo=4)
    assert np.max(np.abs(out[1] - np.float32(0.1))) <= 1e-6 and np.max(np.abs(out[2] + np.float32(0.1))) <= 1e-6


# --------------------------------------------------------------------------
# algo 6 (affine credibility + detail transfer)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("kw", [{}, {"kernel": "jinc"}, {"kernel": "bilinear"}])
@pytest.mark.parametrize("poly", [2, 10])
def test_algo6_sparse_equals_dense(kw, poly):
    # test_backend_matrix.py:104-111 and 132-138 (algo 6 sparse ROI)
    y, u, v = frame(192, 96, bits=32, seed=16, polygons=poly)
    dense, _ = _recon(y, u, v, algo=6, sparse=0, **kw)
    sparse, dbg = _recon(y, u, v, algo=6, sparse=1, **kw)
    assert max(float(np.max(np.abs(dense[p] - sparse[p]))) for p in (1, 2)) <= 1e-6


def test_algo6_flat_luma_keeps_plain():
    # no luma structure: the detail field is zero, algo 6 returns the base
    y = np.full((48, 64), np.float32(0.4))
    rng = np.random.default_rng(17)
    u = (rng.random((24, 32), dtype=np.float32) - 0.5).astype(np.float32)
    v = (rng.random((24, 32), dtype=np.float32) - 0.5).astype(np.float32)
    base, _ = _recon(y, u, v, strength=0.0)
    out, _ = _recon(y, u, v, algo=6, sparse=0)
    assert np.max(np.abs(out[1] - base[1])) <= 1e-7 and np.max(np.abs(out[2] - base[2])) <= 1e-7


def test_algo6_strength_continuity():
    y, u, v = frame(64, 48, bits=32, seed=18, polygons=8)
    zero, _ = _recon(y, u, v, algo=6, strength=0.0, sparse=0)
    tiny, _ = _recon(y, u, v, algo=6, strength=1e-6, sparse=0)
    full, _ = _recon(y, u, v, algo=6, strength=1.0, sparse=0)
    step = max(float(np.max(np.abs(zero[p] - tiny[p]))) for p in (1, 2))
    total = max(float(np.max(np.abs(zero[p] - full[p]))) for p in (1, 2))
    assert total > 1e-4 and step <= total * 3e-6 + 2e-7


# --------------------------------------------------------------------------
# TRecon (temporal taps)
# --------------------------------------------------------------------------

def test_temporal_frame_numbers():
    # plugin.cpp:729-738: clamp to the clip, drop the centre and duplicates
    assert oracle.temporal_frame_numbers(0, 1, 3) == [1]
    assert oracle.temporal_frame_numbers(1, 1, 2) == [0]
    assert oracle.temporal_frame_numbers(5, 9, 2) == [3, 4, 6, 7]
    assert oracle.temporal_frame_numbers(0, 0, 1) == []


def test_trecon_chroma_inconsistent_neighbour_rejected():
    # test_regressions.py:76-93: a neighbour whose chroma disagrees is gated
    # out, so TRecon == single-frame Recon(sparse=0) exactly
    n = 64
    yy, xx = np.mgrid[0:n, 0:n]
    y = (0.35 + 0.08 * np.sin(xx * 0.31) * np.cos(yy * 0.27)).astype(np.float32)
    f0 = (y, np.zeros((32, 32), np.float32), np.zeros((32, 32), np.float32))
    f1 = (y, np.full((32, 32), np.float32(0.2)), np.full((32, 32), np.float32(-0.2)))
    single = oracle.recon_frame(oracle.make_params(n, n, strength=0.8, sparse=0), *f0)
    temporal = oracle.trecon_frame(oracle.trecon_params(strength=0.8, trad=3), f0,
                                   [f1 if k == 1 else f0 for k in oracle.temporal_frame_numbers(0, 1, 3)])
    for p in range(3):
        assert np.array_equal(single[p], temporal[p])


def test_trecon_single_frame_equals_recon():
    y, u, v = frame(96, 64, bits=8, seed=20)
    single = oracle.recon_frame(oracle.make_params(96, 64, sparse=0), y, u, v)
    temporal = oracle.trecon_frame(oracle.trecon_params(), (y, u, v), [])
    for p in range(3):
        assert np.array_equal(single[p], temporal[p])


def test_trecon_444_strength0_identity():
    # test_regressions.py:119-136
    rng = np.random.default_rng(23)
    planes = [rng.uniform(-0.4 if p else 0.0, 0.4 if p else 1.0, (32, 48)).astype(np.float32) for p in range(3)]
    out = oracle.trecon_frame(oracle.trecon_params(strength=0.0), planes, [planes], ss_w=0, ss_h=0)
    for p in range(3):
        assert np.array_equal(out[p], planes[p])


def test_block_match_recovers_shift():
    # a textured frame shifted by (dx, dy) is found by the full search
    rng = np.random.default_rng(21)
    base = rng.random((96, 128), dtype=np.float32)
    for dx, dy in ((3, -2), (0, 5), (-6, 1)):
        nbr = np.roll(base, (dy, dx), axis=(0, 1)).astype(np.float32)
        bw, bh = C.c_int(), C.c_int()
        nb = (128 // 16) * (96 // 16)
        mvx = np.zeros(nb, np.int16)
        mvy = np.zeros(nb, np.int16)
        conf = np.zeros(nb, np.float32)
        L.or_block_match(oracle.fptr(base), oracle.fptr(nbr), 128, 96, 16, 6, 0.02,
                         mvx.ctypes.data_as(C.POINTER(C.c_int16)), mvy.ctypes.data_as(C.POINTER(C.c_int16)),
                         oracle.fptr(conf), C.byref(bw), C.byref(bh))
        inner = np.zeros((6, 8), bool)
        inner[1:-1, 1:-1] = True
        assert np.all(mvx.reshape(6, 8)[inner] == dx) and np.all(mvy.reshape(6, 8)[inner] == dy)
        assert np.all(conf.reshape(6, 8)[inner] > 0.9)


def test_trecon_static_noise_gain():
    # battery.py:236, 313-315: static scene, independent noise per frame ->
    # temporal averaging beats single-frame guided recon
    gy, gu, gv = frame(128, 96, ss_w=0, ss_h=0, bits=32, seed=22, polygons=6, luma_noise=0.0)
    frames = [frame(128, 96, bits=32, seed=22, frame=k, polygons=6, luma_noise=0.0, chroma_noise=0.01)
              for k in range(3)]
    single = oracle.recon_frame(oracle.make_params(128, 96, sparse=0), *frames[1])
    temporal = oracle.trecon_frame(oracle.trecon_params(), frames[1], [frames[0], frames[2]])
    e_s = float(np.mean(np.abs(single[1] - gu)) + np.mean(np.abs(single[2] - gv)))
    e_t = float(np.mean(np.abs(temporal[1] - gu)) + np.mean(np.abs(temporal[2] - gv)))
    assert e_t < e_s, (e_t, e_s)
