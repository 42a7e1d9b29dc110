"""ctypes binding of the C ABI declared in include/lgcr_hip.h.

This is the binding a Python host (tests, benchmark, serving loop) uses; the
VapourSynth plugin glue links the same library from C++.  Nothing here
computes: every call crosses into liblgcr_hip.so, and loading fails loudly
when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import math
import os

import numpy as np

from . import paths

LGCR_OK, LGCR_UNSUPPORTED, LGCR_INVALID_CONTEXT, LGCR_BACKEND_FAILURE = 0, 1, 2, 3
LGCR_OUT_OF_MEMORY, LGCR_REJECTED, LGCR_INVALID_ARGUMENT, LGCR_BUSY = 4, 5, 6, 7
MEMORY_HOST, MEMORY_DEVICE = 0, 1
KERNELS = {"bilinear": 0, "bicubic": 1, "lanczos": 2, "spline16": 3, "spline36": 4, "jinc": 5}
STAGE_COUNT = 19

FETCH = {name: i for i, name in enumerate([
    "y", "u", "v", "jxx", "jxy", "jyy", "lc", "ms", "seed", "mask", "chroma_mask", "indices",
    "plain_u", "plain_v", "out_u", "out_v", "mv", "conf"])}

# exported symbols, in header order
SYMBOLS = [
    "lgcr_abi_version", "lgcr_last_error", "lgcr_params_default", "lgcr_device_open",
    "lgcr_device_close", "lgcr_device_reserved", "lgcr_clip_create", "lgcr_clip_destroy",
    "lgcr_clip_workspace_bytes", "lgcr_recon_frame", "lgcr_recon_batch", "lgcr_clip_set_timing",
    "lgcr_clip_stage_times", "lgcr_clip_fetch", "lgcr_params_default_trecon", "lgcr_trecon_create",
    "lgcr_trecon_frame", "lgcr_sharpen_params_default", "lgcr_sharpen_create", "lgcr_sharpen_frame",
    "lgcr_recon_submit", "lgcr_trecon_submit", "lgcr_ticket_query", "lgcr_ticket_wait",
    "lgcr_clip_set_graphs", "lgcr_clip_set_batching", "lgcr_clip_get_batching", "lgcr_clip_get_tuning", "lgcr_clip_set_tuning", "lgcr_clip_autotune",
    "lgcr_downsample_params_default", "lgcr_downsample_create", "lgcr_downsample_frame",
    "lgcr_ticket_wait_any", "lgcr_downsample_frame_noalloc", "lgcr_check_rcp",
]
STAGE_NAMES = [
    "ConvertInput", "ResampleLuma", "BuildGuideMaps", "BuildTrustMask", "BuildMutualGate",
    "BuildBaseChroma", "BuildLGF", "BuildAffineMaps", "BuildDetail", "ApplyGuidedCorrection",
    "ApplySelector", "ApplyDetailTransfer", "ApplyCollaborativeFilter", "DownsampleBase",
    "DownsampleGuide", "DownsampleCandidateScore", "DownsampleOutput", "BackProject", "ConvertOutput",
]

STAGE_BUILD_BASE_CHROMA = STAGE_NAMES.index("BuildBaseChroma")
STAGE_APPLY_GUIDED_CORRECTION = STAGE_NAMES.index("ApplyGuidedCorrection")


class Params(C.Structure):
    _fields_ = [
        ("kernel", C.c_int), ("kp1", C.c_double), ("kp2", C.c_double), ("support", C.c_double),
        ("radial", C.c_int),
        ("strength", C.c_double), ("sigma", C.c_double), ("sratio", C.c_double), ("sdb", C.c_double),
        ("reg", C.c_double), ("stretch", C.c_double), ("gsigma", C.c_double), ("rescue", C.c_double),
        ("algo", C.c_int), ("ridge", C.c_int), ("cedge", C.c_int),
        ("ms", C.c_double), ("qgate", C.c_double), ("sparse", C.c_int), ("bp", C.c_double),
        ("bm", C.c_int), ("trad", C.c_int), ("tsearch", C.c_int), ("tsad", C.c_double),
        ("ar_margin", C.c_double), ("shift_x", C.c_double), ("shift_y", C.c_double),
        ("loc_set", C.c_int), ("out_w", C.c_int), ("out_h", C.c_int),
    ]


class SharpenParams(C.Structure):
    _fields_ = [("alpha", C.c_double), ("sigma", C.c_double), ("sratio", C.c_double), ("gspatial", C.c_double),
                ("ar_margin", C.c_double)]


class DownsampleParams(C.Structure):
    _fields_ = [("quality", C.c_int), ("kernel", C.c_int), ("strength", C.c_double), ("chroma_location", C.c_int)]


DS_KERNELS = {"spline36": 0, "lanczos3": 1, "binomial": 2}
# loc -> H.273 _ChromaLocation (downsample.cpp:875-893)
DS_LOCATIONS = {"left": 0, "center": 1, "topleft": 2, "top": 3, "bottomleft": 4, "bottom": 5}


class Format(C.Structure):
    _fields_ = [("sample_type", C.c_int), ("bits_per_sample", C.c_int), ("bytes_per_sample", C.c_int),
                ("sub_sampling_w", C.c_int), ("sub_sampling_h", C.c_int), ("width", C.c_int),
                ("height", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("tensor_tile_rows", C.c_int), ("gate_tile_cols", C.c_int)]


class Frame(C.Structure):
    _fields_ = [("data", C.c_void_p * 3), ("stride", C.c_ssize_t * 3), ("memory", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("stage_us", C.c_double * STAGE_COUNT), ("stage_pixels", C.c_uint64 * STAGE_COUNT),
                ("total_us", C.c_double), ("output_pixels", C.c_uint64),
                ("active_output_pixels", C.c_uint64), ("active_chroma_pixels", C.c_uint64),
                ("sparse_total_pixels", C.c_uint64), ("stage_taps", C.c_uint64 * STAGE_COUNT),
                ("taps_visited", C.c_uint64)]


class LgcrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


_lib = None


def lib():
    """Load liblgcr_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(paths.HIP_LIB):
            raise RuntimeError(f"{paths.HIP_LIB} is not built; run __graft_entry__.build()")
        # One HIP runtime per process: torch ships its own libamdhip64.so.7.
        # Loaded first, it satisfies this library's dependency on that soname,
        # so device pointers, streams and events of both live in one runtime;
        # loaded after /opt/rocm's copy, torch would start a second runtime
        # that finds no device.
        if importlib.util.find_spec("torch") is not None:
            import torch  # noqa: F401
        L = C.CDLL(paths.HIP_LIB)
        vp = C.c_void_p
        L.lgcr_abi_version.restype = C.c_int
        L.lgcr_last_error.restype = C.c_char_p
        L.lgcr_params_default.argtypes = [C.POINTER(Params)]
        L.lgcr_device_open.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
        L.lgcr_device_close.argtypes = [vp]
        L.lgcr_device_reserved.argtypes = [vp]
        L.lgcr_device_reserved.restype = C.c_uint64
        L.lgcr_clip_create.argtypes = [vp, C.POINTER(Params), C.POINTER(Format), C.c_int, C.POINTER(vp)]
        L.lgcr_clip_destroy.argtypes = [vp]
        L.lgcr_clip_workspace_bytes.argtypes = [vp]
        L.lgcr_clip_workspace_bytes.restype = C.c_uint64
        L.lgcr_recon_frame.argtypes = [vp, C.POINTER(Frame), C.c_double, C.c_double, C.POINTER(Frame),
                                       C.POINTER(Stats)]
        L.lgcr_recon_batch.argtypes = [vp, C.c_int, C.POINTER(Frame), C.POINTER(Frame)]
        L.lgcr_params_default_trecon.argtypes = [C.POINTER(Params)]
        L.lgcr_trecon_create.argtypes = [vp, C.POINTER(Params), C.POINTER(Format), C.c_int, C.POINTER(vp)]
        L.lgcr_trecon_frame.argtypes = [vp, C.POINTER(Frame), C.POINTER(Frame), C.c_int, C.c_double, C.c_double,
                                        C.POINTER(Frame), C.POINTER(Stats)]
        L.lgcr_clip_set_timing.argtypes = [vp, C.c_int]
        L.lgcr_clip_set_graphs.argtypes = [vp, C.c_int]
        L.lgcr_clip_set_batching.argtypes = [vp, C.c_int]
        L.lgcr_clip_get_batching.argtypes = [vp]
        L.lgcr_clip_get_batching.restype = C.c_int
        L.lgcr_clip_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.lgcr_clip_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.lgcr_clip_autotune.argtypes = [vp, C.POINTER(Frame)]
        L.lgcr_clip_stage_times.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
        L.lgcr_clip_fetch.argtypes = [vp, C.c_int, vp, C.c_size_t, C.POINTER(C.c_size_t)]
        L.lgcr_recon_submit.argtypes = [vp, C.POINTER(Frame), C.c_double, C.c_double, C.POINTER(Frame),
                                        C.POINTER(vp)]
        L.lgcr_trecon_submit.argtypes = [vp, C.POINTER(Frame), C.POINTER(Frame), C.c_int, C.c_double, C.c_double,
                                         C.POINTER(Frame), C.POINTER(vp)]
        L.lgcr_ticket_query.argtypes = [vp]
        L.lgcr_ticket_wait.argtypes = [vp]
        L.lgcr_ticket_wait_any.argtypes = [C.POINTER(vp), C.c_int, C.c_double, C.POINTER(C.c_int)]
        L.lgcr_sharpen_params_default.argtypes = [C.POINTER(SharpenParams)]
        L.lgcr_sharpen_create.argtypes = [vp, C.POINTER(SharpenParams), C.POINTER(Format), C.c_int, C.POINTER(vp)]
        L.lgcr_sharpen_frame.argtypes = [vp, C.POINTER(Frame), C.POINTER(Frame)]
        L.lgcr_downsample_params_default.argtypes = [C.POINTER(DownsampleParams)]
        L.lgcr_downsample_create.argtypes = [vp, C.POINTER(DownsampleParams), C.POINTER(Format), C.c_int,
                                             C.POINTER(vp)]
        L.lgcr_downsample_frame.argtypes = [vp, C.POINTER(Frame), C.POINTER(Frame), C.POINTER(Stats)]
        L.lgcr_downsample_frame_noalloc.argtypes = [vp, C.POINTER(Frame), C.POINTER(Frame)]
        L.lgcr_check_rcp.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def check_rcp(ordinal, lo_bits, hi_bits):
    """(mismatches, first mismatching bit pattern) of the device reciprocal
    against the IEEE quotient over float bit patterns [lo_bits, hi_bits)."""
    n = C.c_uint64(0)
    first = C.c_uint32(0)
    check(lib().lgcr_check_rcp(int(ordinal), int(lo_bits), int(hi_bits), C.byref(n), C.byref(first)))
    return n.value, first.value


def check(status):
    if status != LGCR_OK:
        raise LgcrError(status, lib().lgcr_last_error().decode())


def default_params(**kw) -> Params:
    p = Params()
    lib().lgcr_params_default(C.byref(p))
    kernel = kw.pop("kernel", "lanczos")
    taps = kw.pop("taps", None)
    b = kw.pop("b", None)
    c = kw.pop("c", None)
    p.kernel = KERNELS[kernel]
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
    if "ar" in kw:
        kw["ar_margin"] = kw.pop("ar")
    if "width" in kw:
        kw["out_w"] = kw.pop("width")
    if "height" in kw:
        kw["out_h"] = kw.pop("height")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def make_format(dtype, width, height, ss_w=1, ss_h=1, bits=None) -> Format:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return Format(1, 32, 4, ss_w, ss_h, width, height)
    if bits is None:
        bits = 8 if dt == np.uint8 else 16
    return Format(0, bits, dt.itemsize, ss_w, ss_h, width, height)


def trecon_params(**kw) -> Params:
    """TRecon defaults (plugin.cpp:941-976) with overrides."""
    p = Params()
    lib().lgcr_params_default_trecon(C.byref(p))
    if "ar" in kw:
        kw["ar_margin"] = kw.pop("ar")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def sharpen_params(**kw) -> SharpenParams:
    """Sharpen defaults (plugin.cpp:94-98) with overrides; ar -> ar_margin."""
    p = SharpenParams()
    lib().lgcr_sharpen_params_default(C.byref(p))
    if "ar" in kw:
        kw["ar_margin"] = kw.pop("ar")
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def downsample_params(quality=0, kernel="spline36", strength=1.0, loc="left") -> DownsampleParams:
    """Downsample arguments (downsample.cpp:851-893) as the C struct; kernel
    and loc take the plugin's names."""
    p = DownsampleParams()
    lib().lgcr_downsample_params_default(C.byref(p))
    p.quality = int(quality)
    p.kernel = DS_KERNELS.get(kernel, -1) if isinstance(kernel, str) else int(kernel)
    p.strength = float(strength)
    p.chroma_location = DS_LOCATIONS.get(loc, -1) if isinstance(loc, str) else int(loc)
    return p


def host_frame(planes):
    arrs = [np.ascontiguousarray(a) for a in planes]
    f = Frame((C.c_void_p * 3)(*[a.ctypes.data for a in arrs]), (C.c_ssize_t * 3)(*[a.strides[0] for a in arrs]),
              MEMORY_HOST)
    return f, arrs


class Device:
    def __init__(self, ordinal=0, budget=0):
        h = C.c_void_p()
        check(lib().lgcr_device_open(ordinal, budget, C.byref(h)))
        self.handle = h

    def reserved(self):
        return lib().lgcr_device_reserved(self.handle)

    def close(self):
        if self.handle:
            lib().lgcr_device_close(self.handle)
            self.handle = None


class Clip:
    """One filter instance: parameters + input format, resident on a device."""

    def __init__(self, device: Device, params: Params, fmt: Format, max_inflight=4, temporal=False):
        h = C.c_void_p()
        create = lib().lgcr_trecon_create if temporal else lib().lgcr_clip_create
        check(create(device.handle, C.byref(params), C.byref(fmt), max_inflight, C.byref(h)))
        self.handle = h
        self.fmt = fmt
        self.ow = params.out_w or fmt.width
        self.oh = params.out_h or fmt.height

    def workspace_bytes(self):
        return lib().lgcr_clip_workspace_bytes(self.handle)

    def recon_host(self, y, u, v, shift_x=math.nan, shift_y=math.nan, stats=False, share_luma=False):
        """Host numpy planes in, host numpy planes out (the plugin path)."""
        dt = y.dtype
        out = [None if (share_luma and self.ow == self.fmt.width and self.oh == self.fmt.height)
               else np.empty((self.oh, self.ow), dt)] + [np.empty((self.oh, self.ow), dt) for _ in range(2)]
        ins = [np.ascontiguousarray(a) for a in (y, u, v)]
        fin = Frame((C.c_void_p * 3)(*[a.ctypes.data for a in ins]),
                    (C.c_ssize_t * 3)(*[a.strides[0] for a in ins]), MEMORY_HOST)
        fout = Frame((C.c_void_p * 3)(*[(a.ctypes.data if a is not None else None) for a in out]),
                     (C.c_ssize_t * 3)(*[(a.strides[0] if a is not None else 0) for a in out]), MEMORY_HOST)
        st = Stats()
        check(lib().lgcr_recon_frame(self.handle, C.byref(fin), shift_x, shift_y, C.byref(fout),
                                     C.byref(st) if stats else None))
        if out[0] is None:
            out[0] = y
        return (out, st) if stats else out

    def trecon_host(self, center, nbrs, shift_x=math.nan, shift_y=math.nan):
        """TRecon on host planes: center (y, u, v) and a list of neighbours."""
        fin, keep = host_frame(center)
        n = len(nbrs)
        fn = (Frame * max(1, n))()
        keeps = [keep]
        for k, f in enumerate(nbrs):
            fn[k], kk = host_frame(f)
            keeps.append(kk)
        dt = center[0].dtype
        out = [np.empty((self.oh, self.ow), dt) for _ in range(3)]
        fout, _ = host_frame(out)
        check(lib().lgcr_trecon_frame(self.handle, C.byref(fin), fn, n, shift_x, shift_y, C.byref(fout), None))
        return out

    def recon_device(self, planes_in, planes_out, shift_x=math.nan, shift_y=math.nan):
        """Device pointers: planes_* are sequences of (ptr, stride_bytes)."""
        fin = Frame((C.c_void_p * 3)(*[p for p, _ in planes_in]), (C.c_ssize_t * 3)(*[s for _, s in planes_in]),
                    MEMORY_DEVICE)
        fout = Frame((C.c_void_p * 3)(*[p for p, _ in planes_out]),
                     (C.c_ssize_t * 3)(*[s for _, s in planes_out]), MEMORY_DEVICE)
        check(lib().lgcr_recon_frame(self.handle, C.byref(fin), shift_x, shift_y, C.byref(fout), None))

    def recon_batch(self, frames_in, frames_out):
        """Lists of 3-tuples of (ptr, stride) device planes."""
        n = len(frames_in)
        fin = (Frame * n)()
        fout = (Frame * n)()
        for k in range(n):
            fin[k] = Frame((C.c_void_p * 3)(*[p for p, _ in frames_in[k]]),
                           (C.c_ssize_t * 3)(*[s for _, s in frames_in[k]]), MEMORY_DEVICE)
            fout[k] = Frame((C.c_void_p * 3)(*[p for p, _ in frames_out[k]]),
                            (C.c_ssize_t * 3)(*[s for _, s in frames_out[k]]), MEMORY_DEVICE)
        check(lib().lgcr_recon_batch(self.handle, n, fin, fout))

    def set_timing(self, enable=True):
        check(lib().lgcr_clip_set_timing(self.handle, int(enable)))

    def set_graphs(self, enable=True):
        check(lib().lgcr_clip_set_graphs(self.handle, int(enable)))

    def set_batching(self, frames_per_launch):
        """recon_batch: fused tile pass of up to this many frames per launch (0 = off)."""
        check(lib().lgcr_clip_set_batching(self.handle, int(frames_per_launch)))

    def batching(self):
        return lib().lgcr_clip_get_batching(self.handle)

    def tuning(self):
        t = Tuning()
        check(lib().lgcr_clip_get_tuning(self.handle, C.byref(t)))
        return t.tensor_tile_rows, t.gate_tile_cols

    def set_tuning(self, tensor_tile_rows, gate_tile_cols):
        t = Tuning(tensor_tile_rows, gate_tile_cols)
        check(lib().lgcr_clip_set_tuning(self.handle, C.byref(t)))

    def autotune_host(self, y, u, v):
        """Autotune on one frame of host planes; returns the chosen shapes."""
        f, keep = host_frame((y, u, v))
        check(lib().lgcr_clip_autotune(self.handle, C.byref(f)))
        return self.tuning()

    def autotune_device(self, planes):
        """Autotune on one frame of device planes ((ptr, stride) x 3)."""
        f = Frame((C.c_void_p * 3)(*[p for p, _ in planes]), (C.c_ssize_t * 3)(*[s for _, s in planes]),
                  MEMORY_DEVICE)
        check(lib().lgcr_clip_autotune(self.handle, C.byref(f)))
        return self.tuning()

    def stage_times(self):
        """{stage name: (total_us, launches)} accumulated since set_timing."""
        us = (C.c_double * STAGE_COUNT)()
        n = (C.c_uint64 * STAGE_COUNT)()
        check(lib().lgcr_clip_stage_times(self.handle, us, n))
        return {STAGE_NAMES[i]: (us[i], n[i]) for i in range(STAGE_COUNT) if n[i]}

    def fetch(self, what, dtype=np.float32):
        needed = C.c_size_t()
        check(lib().lgcr_clip_fetch(self.handle, FETCH[what], None, 0, C.byref(needed)))
        buf = np.empty(needed.value // np.dtype(dtype).itemsize, dtype)
        if needed.value:
            check(lib().lgcr_clip_fetch(self.handle, FETCH[what], buf.ctypes.data, needed.value, None))
        return buf

    def close(self):
        if self.handle:
            lib().lgcr_clip_destroy(self.handle)
            self.handle = None


class SharpenClip:
    """One Sharpen instance on a device; frames keep their format."""

    def __init__(self, device: Device, params: SharpenParams, fmt: Format, max_inflight=4):
        h = C.c_void_p()
        check(lib().lgcr_sharpen_create(device.handle, C.byref(params), C.byref(fmt), max_inflight, C.byref(h)))
        self.handle = h
        self.fmt = fmt

    def sharpen_host(self, y, u, v):
        fin, keep = host_frame((y, u, v))
        out = [np.empty_like(a) for a in (y, u, v)]
        fout, _ = host_frame(out)
        check(lib().lgcr_sharpen_frame(self.handle, C.byref(fin), C.byref(fout)))
        return out

    def sharpen_device(self, planes_in, planes_out):
        fin = Frame((C.c_void_p * 3)(*[p for p, _ in planes_in]), (C.c_ssize_t * 3)(*[s for _, s in planes_in]),
                    MEMORY_DEVICE)
        fout = Frame((C.c_void_p * 3)(*[p for p, _ in planes_out]),
                     (C.c_ssize_t * 3)(*[s for _, s in planes_out]), MEMORY_DEVICE)
        check(lib().lgcr_sharpen_frame(self.handle, C.byref(fin), C.byref(fout)))

    def close(self):
        if self.handle:
            lib().lgcr_clip_destroy(self.handle)
            self.handle = None


class DownsampleClip:
    """One Downsample instance: 4:4:4 frames in, 4:2:0 frames out."""

    def __init__(self, device: Device, params: DownsampleParams, fmt: Format, max_inflight=4):
        h = C.c_void_p()
        check(lib().lgcr_downsample_create(device.handle, C.byref(params), C.byref(fmt), max_inflight,
                                           C.byref(h)))
        self.handle = h
        self.fmt = fmt

    def downsample_host(self, y, u, v, stats=False, copy_luma=True):
        """(Y, U, V) 4:4:4 host planes -> (Y, U, V) 4:2:0; with copy_luma
        False the output Y is the input array itself (plane sharing)."""
        fin, keep = host_frame((y, u, v))
        h, w = y.shape
        ou = np.empty((h // 2, w // 2), y.dtype)
        ov = np.empty((h // 2, w // 2), y.dtype)
        oy = np.empty_like(y) if copy_luma else None
        fout = Frame((C.c_void_p * 3)(oy.ctypes.data if copy_luma else None, ou.ctypes.data, ov.ctypes.data),
                     (C.c_ssize_t * 3)(oy.strides[0] if copy_luma else 0, ou.strides[0], ov.strides[0]),
                     MEMORY_HOST)
        st = Stats() if stats else None
        check(lib().lgcr_downsample_frame(self.handle, C.byref(fin), C.byref(fout), C.byref(st) if stats else None))
        out = (oy if copy_luma else y, ou, ov)
        return (out, st) if stats else out

    def downsample_device(self, planes_in, planes_out):
        """Device planes ((ptr, byte stride) x 3); planes_out[0] may be (0, 0)."""
        fin = Frame((C.c_void_p * 3)(*[p for p, _ in planes_in]), (C.c_ssize_t * 3)(*[s for _, s in planes_in]),
                    MEMORY_DEVICE)
        fout = Frame((C.c_void_p * 3)(*[p or None for p, _ in planes_out]),
                     (C.c_ssize_t * 3)(*[s for _, s in planes_out]), MEMORY_DEVICE)
        check(lib().lgcr_downsample_frame(self.handle, C.byref(fin), C.byref(fout), None))

    def downsample_device_noalloc(self, planes_in, planes_out):
        """As downsample_device, through the entry point that allocates
        nothing (a fixed per-GPU scratch array instead of a workspace)."""
        fin = Frame((C.c_void_p * 3)(*[p for p, _ in planes_in]), (C.c_ssize_t * 3)(*[s for _, s in planes_in]),
                    MEMORY_DEVICE)
        fout = Frame((C.c_void_p * 3)(*[p or None for p, _ in planes_out]),
                     (C.c_ssize_t * 3)(*[s for _, s in planes_out]), MEMORY_DEVICE)
        check(lib().lgcr_downsample_frame_noalloc(self.handle, C.byref(fin), C.byref(fout)))

    def close(self):
        if self.handle:
            lib().lgcr_clip_destroy(self.handle)
            self.handle = None
