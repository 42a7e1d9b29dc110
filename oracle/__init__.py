"""LGCR oracle — TEST INFRASTRUCTURE ONLY.

ctypes view of ``liblgcr_oracle.so``, the scalar C restatement of the
reference Recon path (``lgcr_oracle.c``).  Only the tests, the smoke entry
and the CPU-baseline leg of the benchmark load this module; the product
library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblgcr_oracle.so")

KERNELS = {"bilinear": 0, "bicubic": 1, "lanczos": 2, "spline16": 3, "spline36": 4, "jinc": 5}


def build() -> str:
    """Compile the restatement (gcc, no FMA contraction)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


class Params(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("kp1", C.c_double), ("kp2", C.c_double), ("support", C.c_double),
        ("radial", C.c_int), ("lut", C.POINTER(C.c_float)), ("lut_n", C.c_int), ("lut_scale", C.c_float),
        ("strength", C.c_double), ("sigma", C.c_double), ("sratio", C.c_double), ("sdb", C.c_double),
        ("reg", C.c_double), ("stretch", C.c_double), ("gsigma", C.c_double), ("rescue", C.c_double),
        ("algo", C.c_int), ("ridge", C.c_int), ("cedge", C.c_int),
        ("ms", C.c_double), ("qgate", C.c_double), ("sparse", C.c_int), ("bp", C.c_double),
        ("bm", C.c_int), ("ar_margin", C.c_double), ("shift_x", C.c_double), ("shift_y", C.c_double),
        ("out_w", C.c_int), ("out_h", C.c_int),
        ("trad", C.c_int), ("tsearch", C.c_int), ("tsad", C.c_double),
    ]


class Weights(C.Structure):
    _fields_ = [("n", C.c_int), ("sup", C.c_int), ("start", C.POINTER(C.c_int)), ("w", C.POINTER(C.c_float))]


class Axis(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("sup", C.c_int), ("start", C.POINTER(C.c_int)),
        ("w", C.POINTER(C.c_float)), ("am", C.POINTER(C.c_float)),
        ("wsum", C.POINTER(C.c_float)), ("wabs", C.POINTER(C.c_float)),
        ("active", C.POINTER(C.c_int)), ("pos", C.POINTER(C.c_float)), ("lpos", C.POINTER(C.c_float)),
        ("cb_i0", C.POINTER(C.c_int)), ("cb_f", C.POINTER(C.c_float)),
        ("lb_i0", C.POINTER(C.c_int)), ("lb_f", C.POINTER(C.c_float)),
    ]


class Workset(C.Structure):
    _fields_ = [
        ("mask_w", C.c_int), ("mask_h", C.c_int), ("out_w", C.c_int), ("out_h", C.c_int),
        ("chroma_w", C.c_int), ("chroma_h", C.c_int),
        ("mask", C.POINTER(C.c_uint8)), ("indices", C.POINTER(C.c_uint32)), ("n_indices", C.c_size_t),
        ("index_row_offsets", C.POINTER(C.c_size_t)), ("spans", C.POINTER(C.c_int)), ("n_spans", C.c_size_t),
        ("span_row_offsets", C.POINTER(C.c_size_t)), ("chroma_mask", C.POINTER(C.c_uint8)),
        ("active_chroma", C.c_size_t),
    ]


class FrameFormat(C.Structure):
    _fields_ = [("sample_type", C.c_int), ("bits", C.c_int), ("bytes_per_sample", C.c_int),
                ("ss_w", C.c_int), ("ss_h", C.c_int), ("width", C.c_int), ("height", C.c_int)]


class Debug(C.Structure):
    _fields_ = [
        ("sw", C.c_int), ("sh", C.c_int), ("cw", C.c_int), ("ch", C.c_int), ("ow", C.c_int), ("oh", C.c_int),
        ("y", C.POINTER(C.c_float)), ("u", C.POINTER(C.c_float)), ("v", C.POINTER(C.c_float)),
        ("jxx", C.POINTER(C.c_float)), ("jxy", C.POINTER(C.c_float)), ("jyy", C.POINTER(C.c_float)),
        ("lc", C.POINTER(C.c_float)), ("ms", C.POINTER(C.c_float)),
        ("seed", C.POINTER(C.c_uint8)), ("mask", C.POINTER(C.c_uint8)),
        ("ws", Workset), ("has_ws", C.c_int),
        ("plain_u", C.POINTER(C.c_float)), ("plain_v", C.POINTER(C.c_float)),
        ("out_u", C.POINTER(C.c_float)), ("out_v", C.POINTER(C.c_float)),
    ]


class Guide(C.Structure):
    _fields_ = [
        ("y", C.POINTER(C.c_float)), ("sw", C.c_int), ("sh", C.c_int),
        ("jxx", C.POINTER(C.c_float)), ("jxy", C.POINTER(C.c_float)), ("jyy", C.POINTER(C.c_float)),
        ("lc", C.POINTER(C.c_float)), ("ms", C.POINTER(C.c_float)),
        ("nbrs", C.c_void_p), ("n_nbrs", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        u8p = C.POINTER(C.c_uint8)
        L.or_kernel_eval.restype = C.c_double
        L.or_kernel_eval.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.or_jinc_windowed.restype = C.c_double
        L.or_jinc_windowed.argtypes = [C.c_double, C.c_double]
        L.or_jinc_lut.argtypes = [C.c_double, fp, C.c_int]
        L.or_lut_lookup.restype = C.c_float
        L.or_lut_lookup.argtypes = [fp, C.c_int, C.c_float]
        L.or_build_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.POINTER(Weights)]
        L.or_free_weights.argtypes = [C.POINTER(Weights)]
        L.or_build_axis.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(Params), C.POINTER(Axis)]
        L.or_free_axis.argtypes = [C.POINTER(Axis)]
        L.or_sobel_tensor.argtypes = [fp, C.c_int, C.c_int, fp, fp, fp, u8p, C.c_float]
        L.or_lc_map.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_double, fp]
        L.or_dilate.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p]
        L.or_build_workset.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Axis),
                                       C.POINTER(Axis), C.c_int, C.c_int, C.POINTER(Workset)]
        L.or_free_workset.argtypes = [C.POINTER(Workset)]
        L.or_workset_output_dense.argtypes = [C.POINTER(Workset)]
        L.or_workset_chroma_dense.argtypes = [C.POINTER(Workset)]
        L.or_mutual_gate.argtypes = [fp, fp, fp, C.c_int, C.c_int, u8p, C.c_int, fp]
        L.or_plain_chroma.argtypes = [C.POINTER(Params), fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, fp, fp]
        L.or_guided.argtypes = [C.POINTER(Params), fp, fp, C.c_int, C.c_int, C.POINTER(Guide),
                                C.POINTER(Workset), fp, fp]
        L.or_lgf_maps.argtypes = [fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, fp, fp, C.c_double, C.c_int, u8p, fp]
        L.or_back_project.argtypes = [fp, C.c_int, C.c_int, fp, C.c_int, C.c_int, C.c_float]
        L.or_sharpen_frame.argtypes = [C.POINTER(SharpenParams), C.POINTER(FrameFormat), C.c_void_p * 3,
                                       C.c_ssize_t * 3, C.c_void_p * 3, C.c_ssize_t * 3]
        L.or_sharpen_plane.argtypes = [fp, C.c_int, C.c_int, fp, fp, C.c_int, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.POINTER(SharpenParams)]
        L.or_collaborative.argtypes = [fp, C.c_int, C.c_int, fp, fp, C.c_int, C.c_int, C.c_float]
        L.or_recon_frame.argtypes = [C.POINTER(Params), C.POINTER(FrameFormat), C.c_void_p * 3,
                                     C.c_ssize_t * 3, C.c_void_p * 3, C.c_ssize_t * 3]
        L.or_last_debug.restype = C.POINTER(Debug)
        L.or_trecon_frame.argtypes = [C.POINTER(Params), C.POINTER(FrameFormat), C.c_void_p * 3, C.c_ssize_t * 3,
                                      C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_ssize_t),
                                      C.c_void_p * 3, C.c_ssize_t * 3]
        L.or_affine_debug.argtypes = [fp]
        L.or_lc_map.restype = None
        i16p = C.POINTER(C.c_int16)
        L.or_block_match.argtypes = [fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, i16p, i16p, fp,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.or_chroma_consistency.argtypes = [fp, fp, fp, fp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, i16p, i16p,
                                            C.c_int, C.c_int, fp]
        _lib = L
    return _lib


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def u8ptr(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


# ---------------------------------------------------------------------------
# parameter handling (defaults of plugin.cpp:1025-1133)
# ---------------------------------------------------------------------------

def make_params(out_w, out_h, kernel="lanczos", taps=None, b=None, c=None, strength=0.8,
                sigma=0.01, sratio=0.15, sdb=3.0, stretch=1.0, gsigma=2.5, rescue=1.0,
                ridge=1, cedge=0, ar=0.0, reg=0.005, shift_x=-0.5, shift_y=0.0, sparse=1,
                bp=0.0, ms=1.0, qgate=1.0, bm=0, algo=2, trad=1, tsearch=6, tsad=0.02):
    p = Params()
    k = KERNELS[kernel]
    p.kernel = k
    p.kp1, p.kp2 = 3.0, 0.6
    if kernel == "bilinear":
        p.support = 1.0
    elif kernel == "bicubic":
        p.support = 2.0
        p.kp1 = 0.0 if b is None else float(b)
        p.kp2 = 0.6 if c is None else float(c)
    elif kernel in ("lanczos", "jinc"):
        t = 3 if taps is None else int(taps)
        p.kp1 = float(t)
        p.support = float(t)
    elif kernel == "spline16":
        p.support = 2.0
    else:
        p.support = 3.0
    p.radial = 1 if kernel == "jinc" else 0
    p._lut_keep = None
    if p.radial:
        n = int(p.support * 1024) + 2
        lut = np.zeros(n, np.float32)
        lib().or_jinc_lut(p.support, fptr(lut), n)
        p._lut_keep = lut
        p.lut = fptr(lut)
        p.lut_n = n
        p.lut_scale = 1024.0
    p.strength, p.sigma, p.sratio, p.sdb = strength, sigma, sratio, sdb
    p.reg, p.stretch, p.gsigma, p.rescue = reg, stretch, gsigma, rescue
    p.algo, p.ridge, p.cedge = algo, ridge, cedge
    p.ms, p.qgate, p.sparse, p.bp, p.bm = ms, qgate, sparse, bp, bm
    p.ar_margin, p.shift_x, p.shift_y = ar, shift_x, shift_y
    p.out_w, p.out_h = out_w, out_h
    p.trad, p.tsearch, p.tsad = trad, tsearch, tsad
    return p


def recon_frame(params: Params, y, u, v, bits=None, ss_w=1, ss_h=1):
    """Run the whole Recon frame on host planes (numpy uint8/uint16/float32).

    Returns (out_y, out_u, out_v) as arrays of the same sample type.
    """
    is_float = y.dtype == np.float32
    if bits is None:
        bits = 32 if is_float else (8 if y.dtype == np.uint8 else 16)
    fmt = FrameFormat(1 if is_float else 0, bits, y.dtype.itemsize, ss_w, ss_h, y.shape[1], y.shape[0])
    oh, ow = params.out_h, params.out_w
    outs = [np.zeros((oh, ow), y.dtype) for _ in range(3)]
    ins = [np.ascontiguousarray(a) for a in (y, u, v)]
    pin = (C.c_void_p * 3)(*[a.ctypes.data for a in ins])
    sin = (C.c_ssize_t * 3)(*[a.strides[0] for a in ins])
    pout = (C.c_void_p * 3)(*[a.ctypes.data for a in outs])
    sout = (C.c_ssize_t * 3)(*[a.strides[0] for a in outs])
    rc = lib().or_recon_frame(C.byref(params), C.byref(fmt), pin, sin, pout, sout)
    if rc != 0:
        raise RuntimeError(f"oracle recon_frame failed ({rc})")
    return outs


class SharpenParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("sigma", C.c_double), ("sratio", C.c_double), ("gspatial", C.c_double),
                ("ar_margin", C.c_double)]


def sharpen_params(alpha=0.3, sigma=0.01, sratio=0.15, gspatial=1.2, ar=0.0):
    """Sharpen defaults (plugin.cpp:94-98)."""
    return SharpenParams(alpha, sigma, sratio, gspatial, ar)


def sharpen_frame(params, y, u, v, bits=None, ss_w=1, ss_h=1):
    """Whole Sharpen frame (sharpenGetFrame, plugin.cpp:14-90) on host planes."""
    is_float = y.dtype == np.float32
    if bits is None:
        bits = 32 if is_float else (8 if y.dtype == np.uint8 else 16)
    fmt = FrameFormat(1 if is_float else 0, bits, y.dtype.itemsize, ss_w, ss_h, y.shape[1], y.shape[0])
    ins = [np.ascontiguousarray(a) for a in (y, u, v)]
    outs = [np.zeros_like(a) for a in ins]
    rc = lib().or_sharpen_frame(C.byref(params), C.byref(fmt), (C.c_void_p * 3)(*[a.ctypes.data for a in ins]),
                                (C.c_ssize_t * 3)(*[a.strides[0] for a in ins]),
                                (C.c_void_p * 3)(*[a.ctypes.data for a in outs]),
                                (C.c_ssize_t * 3)(*[a.strides[0] for a in outs]))
    if rc != 0:
        raise RuntimeError(f"oracle sharpen_frame failed ({rc})")
    return outs


def sharpen_plane(params, p, lc=None, gy=None, rw=1.0, rh=1.0, sx=0.0, sy=0.0):
    """sharpenPlane on one float plane; returns the sharpened copy."""
    out = np.array(p, np.float32, copy=True)
    h, w = out.shape
    g = np.ascontiguousarray(gy if gy is not None else out, np.float32)
    lcp = fptr(np.ascontiguousarray(lc, np.float32)) if lc is not None else None
    lib().or_sharpen_plane(fptr(out), w, h, lcp, fptr(g), g.shape[1], g.shape[0], rw, rh, sx, sy, C.byref(params))
    return out


def collaborative(guide, u, v, sigma):
    """Collaborative chroma filter (bm.cpp:250-442) on float planes; returns
    filtered copies of (u, v).  guide may differ in size (bilinear resize)."""
    g = np.ascontiguousarray(guide, np.float32)
    uo = np.array(u, np.float32, copy=True)
    vo = np.array(v, np.float32, copy=True)
    h, w = uo.shape
    lib().or_collaborative(fptr(g), g.shape[1], g.shape[0], fptr(uo), fptr(vo), w, h, float(sigma))
    return uo, vo


def last_debug():
    """Copies of the intermediates of the last recon_frame call."""
    d = lib().or_last_debug().contents

    def arr(ptr, n, dt):
        if not ptr:
            return None
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

    ns, nc, no = d.sw * d.sh, d.cw * d.ch, d.ow * d.oh
    out = {
        "y": arr(d.y, ns, np.float32), "u": arr(d.u, nc, np.float32), "v": arr(d.v, nc, np.float32),
        "jxx": arr(d.jxx, ns, np.float32), "jxy": arr(d.jxy, ns, np.float32), "jyy": arr(d.jyy, ns, np.float32),
        "lc": arr(d.lc, nc, np.float32), "ms": arr(d.ms, nc, np.float32),
        "seed": arr(d.seed, ns, np.uint8), "mask": arr(d.mask, ns, np.uint8),
        "plain_u": arr(d.plain_u, no, np.float32), "plain_v": arr(d.plain_v, no, np.float32),
        "out_u": arr(d.out_u, no, np.float32), "out_v": arr(d.out_v, no, np.float32),
        "dims": (d.sw, d.sh, d.cw, d.ch, d.ow, d.oh),
    }
    if d.has_ws:
        ws = d.ws
        out["indices"] = arr(ws.indices, ws.n_indices, np.uint32) if ws.n_indices else np.zeros(0, np.uint32)
        out["chroma_mask"] = arr(ws.chroma_mask, nc, np.uint8)
        out["active_chroma"] = ws.active_chroma
    return out


def temporal_frame_numbers(n, n_last, trad):
    """Neighbour frames of TRecon: n-trad..n+trad clamped to the clip,
    centre and duplicates removed (plugin.cpp:729-738)."""
    out = []
    for k in range(-trad, trad + 1):
        fn = min(max(n + k, 0), n_last)
        if fn != n and fn not in out:
            out.append(fn)
    return out


def trecon_params(strength=0.8, sigma=0.01, sratio=0.15, sdb=3.0, gsigma=2.5, stretch=1.0, ar=0.0,
                  ridge=1, sparse=0, ms=1.0, trad=1, tsearch=6, tsad=0.02, bm=0, shift_x=-0.5, shift_y=0.0):
    """TRecon parameters (plugin.cpp:939-976): lanczos 3, algo 2."""
    return make_params(0, 0, kernel="lanczos", taps=3, strength=strength, sigma=sigma, sratio=sratio, sdb=sdb,
                       gsigma=gsigma, stretch=stretch, ar=ar, ridge=ridge, sparse=sparse, ms=ms, trad=trad,
                       tsearch=tsearch, tsad=tsad, bm=bm, shift_x=shift_x, shift_y=shift_y)


def trecon_frame(params, center, nbrs, bits=None, ss_w=1, ss_h=1):
    """center: (y, u, v); nbrs: list of (y, u, v) neighbour frames."""
    y, u, v = center
    is_float = y.dtype == np.float32
    if bits is None:
        bits = 32 if is_float else (8 if y.dtype == np.uint8 else 16)
    fmt = FrameFormat(1 if is_float else 0, bits, y.dtype.itemsize, ss_w, ss_h, y.shape[1], y.shape[0])
    h, w = y.shape
    outs = [np.zeros((h, w), y.dtype) for _ in range(3)]
    ins = [np.ascontiguousarray(a) for a in center]
    nb = [np.ascontiguousarray(a) for f in nbrs for a in f]
    nptr = (C.c_void_p * max(1, len(nb)))(*[a.ctypes.data for a in nb])
    nstr = (C.c_ssize_t * max(1, len(nb)))(*[a.strides[0] for a in nb])
    rc = lib().or_trecon_frame(C.byref(params), C.byref(fmt), (C.c_void_p * 3)(*[a.ctypes.data for a in ins]),
                               (C.c_ssize_t * 3)(*[a.strides[0] for a in ins]), len(nbrs), nptr, nstr,
                               (C.c_void_p * 3)(*[a.ctypes.data for a in outs]),
                               (C.c_ssize_t * 3)(*[a.strides[0] for a in outs]))
    if rc != 0:
        raise RuntimeError(f"oracle trecon_frame failed ({rc})")
    return outs


class DownsampleParams(C.Structure):
    _fields_ = [("quality", C.c_int), ("kernel", C.c_int), ("strength", C.c_float), ("shift_x", C.c_float),
                ("shift_y", C.c_float)]


DOWNSAMPLE_KERNELS = {"spline36": 0, "lanczos3": 1, "binomial": 2}
# loc -> (H.273 _ChromaLocation, shift x, shift y) (downsample.cpp:875-893)
DOWNSAMPLE_LOCATIONS = {"left": (0, -0.5, 0.0), "center": (1, 0.0, 0.0), "topleft": (2, -0.5, -0.5),
                        "top": (3, 0.0, -0.5), "bottomleft": (4, -0.5, 0.5), "bottom": (5, 0.0, 0.5)}


def downsample_params(quality=0, kernel="spline36", strength=1.0, loc="left"):
    _, sx, sy = DOWNSAMPLE_LOCATIONS[loc]
    return DownsampleParams(int(quality), DOWNSAMPLE_KERNELS[kernel], float(strength), sx, sy)


def _ds_call(fn, params, y, u, v, bits, outs):
    is_float = y.dtype == np.float32
    if bits is None:
        bits = 32 if is_float else (8 if y.dtype == np.uint8 else 16)
    fmt = FrameFormat(1 if is_float else 0, bits, y.dtype.itemsize, 0, 0, y.shape[1], y.shape[0])
    ins = [np.ascontiguousarray(a) for a in (y, u, v)]
    return fmt, ins, fn(C.byref(params), C.byref(fmt), (C.c_void_p * 3)(*[a.ctypes.data for a in ins]),
                        (C.c_ssize_t * 3)(*[a.strides[0] for a in ins]), *outs)


def downsample_frame(params, y, u, v, bits=None):
    """Downsample 4:4:4 -> 4:2:0 (downsampleGetFrame, downsample.cpp:725-765):
    returns (Y, U, V) with Y the source plane and U, V at half size in the
    source sample type."""
    h, w = y.shape
    ou = np.zeros((h // 2, w // 2), y.dtype)
    ov = np.zeros((h // 2, w // 2), y.dtype)
    L = lib()
    L.or_downsample_frame.restype = C.c_int
    outs = ((C.c_void_p * 3)(None, ou.ctypes.data, ov.ctypes.data), (C.c_ssize_t * 3)(0, ou.strides[0], ov.strides[0]))
    _, _, rc = _ds_call(L.or_downsample_frame, params, y, u, v, bits, outs)
    if rc != 0:
        raise RuntimeError(f"oracle downsample_frame failed ({rc})")
    return y, ou, ov


def downsample_base(params, y, u, v, bits=None):
    """The plain separable baseline (BaselineRing) as float U, V planes."""
    h, w = y.shape
    bu = np.zeros((h // 2, w // 2), np.float32)
    bv = np.zeros((h // 2, w // 2), np.float32)
    L = lib()
    L.or_downsample_base.restype = C.c_int
    _, _, rc = _ds_call(L.or_downsample_base, params, y, u, v, bits, (fptr(bu), fptr(bv)))
    if rc != 0:
        raise RuntimeError(f"oracle downsample_base failed ({rc})")
    return bu, bv
