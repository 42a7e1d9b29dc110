"""Host-side mirror of the LGCR VapourSynth plugin surface (src/plugin.cpp)
over the MI355X backend.

``core.lgcr.Recon``, ``core.lgcr.TRecon``, ``core.lgcr.Sharpen`` and
``core.lgcr.Downsample`` take the reference's arguments with the same names,
defaults, ranges and error messages (lgcrCreate, plugin.cpp:1012-1158;
tReconCreate, plugin.cpp:921-990; sharpenCreate, plugin.cpp:85-111;
downsampleCreate, downsample.cpp:816-906) and return clips whose frames are
computed by ``liblgcr_hip.so``.  Clips and frames are small in-memory
stand-ins for VapourSynth's (VapourSynth itself is not a dependency):
planar numpy planes plus a property dict.  Per frame the filters apply the
reference's property rules: H.273 ``_ChromaLocation`` siting on subsampled
axes unless ``loc`` is given (plugin.cpp:157-175), the prop removed from 4:4:4
outputs (plugin.cpp:268-271, 840-841), the source luma plane shared by
same-size Recon (plugin.cpp:265-269), ``_LGCR_*`` profile props when
``LGCR_PROFILE`` is set (plugin.cpp:114-155).
"""
from __future__ import annotations

import contextlib
import math
import os
import threading
from dataclasses import dataclass, field, replace

import numpy as np

from . import abi

YUV = "YUV"
INTEGER, FLOAT = 0, 1


class Error(Exception):
    """Filter creation or frame errors (vs.Error in VapourSynth)."""


@dataclass(frozen=True)
class VideoFormat:
    color_family: str
    sample_type: int
    bits_per_sample: int
    subsampling_w: int
    subsampling_h: int

    @property
    def bytes_per_sample(self):
        return 4 if self.sample_type == FLOAT else (1 if self.bits_per_sample <= 8 else 2)

    @property
    def dtype(self):
        return np.float32 if self.sample_type == FLOAT else (np.uint8 if self.bytes_per_sample == 1 else np.uint16)


def query_video_format(color_family, sample_type, bits, ss_w, ss_h):
    return VideoFormat(color_family, sample_type, bits, ss_w, ss_h)


@dataclass
class VideoFrame:
    format: VideoFormat
    planes: list
    props: dict = field(default_factory=dict)

    @property
    def width(self):
        return self.planes[0].shape[1]

    @property
    def height(self):
        return self.planes[0].shape[0]

    def __getitem__(self, p):
        return self.planes[p]


@contextlib.contextmanager
def _frame_error(name):
    """A backend failure is the filter's error, prefixed with its name: while
    producing a frame that frame's error (vsapi->setFilterError), leaving the
    clip and the process usable; while creating the filter a creation error
    (vsapi->mapSetError)."""
    try:
        yield
    except abi.LgcrError as e:
        raise Error(f"{name}: {e}") from e


class VideoNode:
    """A clip: constant format and size, frames produced on request."""

    def __init__(self, fmt, width, height, num_frames, get):
        self.format, self.width, self.height, self.num_frames = fmt, width, height, num_frames
        self._get = get

    def get_frame(self, n):
        if not 0 <= n < self.num_frames:
            raise Error(f"requested frame {n} is out of range")
        return self._get(n)


def clip_from_planes(frames, fmt: VideoFormat, props=None):
    """Source clip over host planes: frames is a list of (Y, U, V) arrays;
    props a dict or a list of per-frame dicts."""
    h, w = frames[0][0].shape

    def get(n):
        p = props[n] if isinstance(props, list) else (props or {})
        return VideoFrame(fmt, list(frames[n]), dict(p))

    return VideoNode(fmt, w, h, len(frames), get)


# ---------------------------------------------------------------------------
# argument handling shared by the filters
# ---------------------------------------------------------------------------

def _check_args(fn, given, allowed):
    for k in given:
        if k not in allowed:
            raise Error(f"{fn}: argument {k} is not expected")


def _get(args, key, default):
    v = args.get(key)
    return default if v is None else v


def _getf(args, key, default):
    """A float argument read like the reference reads it: a given value
    narrowed to float (mapGetFloatSaturated), the double default otherwise."""
    v = args.get(key)
    return float(default) if v is None else float(np.float32(v))


def _device():
    with _Device.lock:
        if _Device.dev is None:
            _Device.dev = abi.Device(0)
        return _Device.dev


class _Device:
    lock = threading.Lock()
    dev = None


def _profiling():
    v = os.environ.get("LGCR_PROFILE", "")
    return bool(v) and v[0] != "0"


def _siting(props, fmt, shift_x, shift_y, loc_set):
    """applyFrameChromaSiting (plugin.cpp:157-175)."""
    if fmt.subsampling_w == 0:
        shift_x = 0.0
    if fmt.subsampling_h == 0:
        shift_y = 0.0
    if loc_set or "_ChromaLocation" not in props:
        return shift_x, shift_y
    cl = int(props["_ChromaLocation"])
    if fmt.subsampling_w > 0:
        shift_x = -0.5 if cl in (0, 2, 4) else 0.0
    if fmt.subsampling_h > 0:
        shift_y = -0.5 if cl in (2, 3) else (0.5 if cl in (4, 5) else 0.0)
    return shift_x, shift_y


def _abi_format(fmt, w, h):
    return abi.Format(fmt.sample_type, fmt.bits_per_sample, fmt.bytes_per_sample, fmt.subsampling_w,
                      fmt.subsampling_h, w, h)


# frame-property names of the stages (stageName, src/pipeline.h:222-244), in
# the backend's LGCR_STAGE_* order
PROP_STAGE_NAMES = ["input_conversion", "luma_resample", "buildGuideMaps", "buildTrustMask", "buildMutualGate",
                    "plainChroma", "buildLGF", "buildAffineMaps", "buildDetail", "reconstructChroma",
                    "selectorBlend", "detailTransfer", "collaborative_chroma", "downsample_base",
                    "downsample_guide", "downsample_candidate_score", "downsample_output", "backProject",
                    "output_conversion"]


# CPU sub-stage timer names (cpuProfileSlotName, src/pipeline.h:282-312)
CPU_SLOTS = ["guide_sobel", "guide_tensor", "guide_lcMap", "affine_candidate_box", "affine_candidate_bilinear",
             "affine_candidate_bicubic", "affine_minmax_u", "affine_minmax_v", "affine_window_moments",
             "affine_finalize_median", "guided_active_rows", "guided_metadata", "guided_tap_accumulation",
             "guided_normalization_selector", "trust_seed", "trust_dilate", "mutual_gradients", "mutual_gate",
             "plain_horizontal", "plain_vertical", "lgf_moments", "lgf_finalize", "affine_rolling",
             "affine_consume", "detail_reconstruct", "detail_nyquist", "detail_transfer"]


def _profile_props(props, st, fmt, sw, sh, ow, oh, bm):
    """writeProfile (plugin.cpp:126-155) from the backend's per-frame stats.
    Conversion work follows the reference's counting (plugin.cpp:250-253,
    708-710): float input planes are viewed, not converted, and float output
    is converted only with bm."""
    is_float = fmt.sample_type == FLOAT
    cw, ch = sw >> fmt.subsampling_w, sh >> fmt.subsampling_h
    pixels = [int(v) for v in st.stage_pixels]
    pixels[0] = 0 if is_float else sw * sh + 2 * cw * ch
    pixels[18] = 2 * ow * oh if (not is_float or bm) else 0
    for i, name in enumerate(PROP_STAGE_NAMES):
        props[f"_LGCR_{name}_us"] = float(st.stage_us[i])
        props[f"_LGCR_{name}_pixels"] = pixels[i]
        props[f"_LGCR_{name}_taps"] = int(st.stage_taps[i])
    for slot in CPU_SLOTS:
        # CPU time of the reference's sub-stage timers: the backend computes
        # on the device, so none is spent on the host
        props[f"_LGCR_cpu_{slot}_us"] = 0.0
    props["_LGCR_total_us"] = float(st.total_us)
    props["_LGCR_output_pixels"] = int(st.output_pixels)
    props["_LGCR_taps_visited"] = int(st.taps_visited)
    props["_LGCR_sparse_active_ratio"] = (float(st.active_output_pixels) / st.sparse_total_pixels
                                          if st.sparse_total_pixels else 1.0)


# ---------------------------------------------------------------------------
# Recon (lgcrCreate / lgcrGetFrame)
# ---------------------------------------------------------------------------

RECON_ARGS = ("width", "height", "kernel", "taps", "algo", "b", "c", "strength", "sigma", "sratio", "sdb",
              "stretch", "gsigma", "rescue", "ridge", "cedge", "ar", "reg", "loc", "sparse", "bp", "ms", "qgate",
              "bm")


def recon_params(clip: VideoNode, **args):
    """lgcrCreate's argument parsing and validation (plugin.cpp:1012-1142);
    returns the backend parameters."""
    _check_args("Recon", args, RECON_ARGS)
    fmt = clip.format
    if fmt.color_family != YUV or (fmt.sample_type == FLOAT and fmt.bits_per_sample != 32):
        raise Error("LGCR: input must be planar YUV (8-16 bit int or 32 bit float)")
    if fmt.subsampling_w > 1 or fmt.subsampling_h > 1:
        raise Error("LGCR: only 4:4:4 / 4:2:2 / 4:2:0 input is supported")
    if clip.width == 0 or clip.height == 0:
        raise Error("LGCR: constant input dimensions required")
    p = abi.Params()
    abi.lib().lgcr_params_default(p)
    p.out_w = int(_get(args, "width", clip.width))
    p.out_h = int(_get(args, "height", clip.height))
    if p.out_w <= 0 or p.out_h <= 0:
        raise Error("LGCR: invalid width/height")
    k = _get(args, "kernel", "lanczos")
    if isinstance(k, bytes):
        k = k.decode()
    p.radial = 0
    if k == "bilinear":
        p.kernel, p.support = abi.KERNELS["bilinear"], 1.0
    elif k == "bicubic":
        p.kernel, p.support = abi.KERNELS["bicubic"], 2.0
        p.kp1 = _getf(args, "b", 0.0)
        p.kp2 = _getf(args, "c", 0.6)
    elif k in ("lanczos", "jinc"):
        taps = int(_get(args, "taps", 3))
        if taps < 1 or taps > 64:
            raise Error(f"LGCR: {k} taps must be in the range 1..64")
        p.kernel = abi.KERNELS[k]
        p.kp1 = float(taps)
        p.support = float(taps)
        p.radial = 1 if k == "jinc" else 0
    elif k == "spline16":
        p.kernel, p.support = abi.KERNELS["spline16"], 2.0
    elif k == "spline36":
        p.kernel, p.support = abi.KERNELS["spline36"], 3.0
    else:
        raise Error('LGCR: kernel must be "bilinear", "bicubic", "lanczos", "spline16", "spline36" or "jinc"')
    p.strength = _getf(args, "strength", 0.8)
    p.sigma = _getf(args, "sigma", 0.01)
    p.sratio = _getf(args, "sratio", 0.15)
    p.algo = int(_get(args, "algo", 2))
    if p.algo not in (2, 4, 6):
        raise Error("LGCR: algo must be 2 (guided), 4 (selector) or 6 (detail transfer)")
    p.sdb = _getf(args, "sdb", 3.0)
    p.stretch = _getf(args, "stretch", 1.0)
    p.gsigma = _getf(args, "gsigma", 2.5)
    p.rescue = _getf(args, "rescue", 1.0)
    p.ridge = 1 if int(_get(args, "ridge", 1)) != 0 else 0
    p.cedge = 1 if int(_get(args, "cedge", 0)) != 0 else 0
    p.ms = _getf(args, "ms", 1.0)
    p.qgate = _getf(args, "qgate", 1.0)
    p.ar_margin = _getf(args, "ar", 0.0)
    p.reg = _getf(args, "reg", 0.005)
    loc = args.get("loc")
    p.shift_x, p.shift_y, p.loc_set = -0.5, 0.0, 0
    if loc is not None:
        if isinstance(loc, bytes):
            loc = loc.decode()
        if loc == "left":
            p.shift_x = -0.5
        elif loc == "center":
            p.shift_x = 0.0
        else:
            raise Error('LGCR: loc must be "left" or "center"')
        p.loc_set = 1
    p.sparse = 1 if int(_get(args, "sparse", 1)) != 0 else 0
    p.bm = 1 if int(_get(args, "bm", 0)) != 0 else 0
    p.bp = _getf(args, "bp", 0.0)
    if fmt.subsampling_w == 0:
        p.shift_x = 0.0
    if fmt.subsampling_h == 0:
        p.shift_y = 0.0
    if (p.strength < 0.0 or p.strength > 1.0 or p.sigma <= 0.0 or p.sratio <= 0.0 or p.sdb <= 0.0
            or p.stretch < 0.0 or p.gsigma <= 0.0 or p.reg <= 0.0 or p.bp < 0.0 or p.bp > 1.0 or p.ms < 0.0
            or p.ms > 1.0 or p.qgate < 0.0 or p.qgate > 1.0 or p.rescue < 0.0 or p.rescue > 1.0):
        raise Error("LGCR: invalid parameter range (need 0<=strength/rescue<=1, 0<=bp/ms/qgate<=1, "
                    "sigma/sratio/sdb/gsigma/reg>0, stretch>=0)")
    return p


def Recon(clip: VideoNode, **args) -> VideoNode:
    p = recon_params(clip, **args)
    fmt_in = clip.format
    fmt_out = replace(fmt_in, subsampling_w=0, subsampling_h=0)
    with _frame_error("LGCR"):
        backend = abi.Clip(_device(), p, _abi_format(fmt_in, clip.width, clip.height))
    ow, oh = p.out_w, p.out_h
    same = ow == clip.width and oh == clip.height
    shift_x, shift_y, loc_set = p.shift_x, p.shift_y, bool(p.loc_set)

    def get(n):
        src = clip.get_frame(n)
        sx, sy = _siting(src.props, fmt_in, shift_x, shift_y, loc_set)
        profile = _profiling()
        with _frame_error("LGCR"):
            res = backend.recon_host(*src.planes, shift_x=sx, shift_y=sy, stats=profile, share_luma=same)
        out, st = res if profile else (res, None)
        props = dict(src.props)
        props.pop("_ChromaLocation", None)
        if profile:
            _profile_props(props, st, fmt_in, clip.width, clip.height, ow, oh, bool(p.bm))
        return VideoFrame(fmt_out, [src.planes[0] if same else out[0], out[1], out[2]], props)

    node = VideoNode(fmt_out, ow, oh, clip.num_frames, get)
    node._backend = backend
    return node


# ---------------------------------------------------------------------------
# TRecon (tReconCreate / tReconGetFrame)
# ---------------------------------------------------------------------------

TRECON_ARGS = ("strength", "sigma", "sratio", "sdb", "gsigma", "stretch", "ar", "ridge", "sparse", "ms", "trad",
               "tsearch", "tsad", "bm")


def temporal_frame_numbers(n, n_last, trad):
    """plugin.cpp:729-738: n-trad..n+trad clamped, centre and duplicates
    removed."""
    out = []
    for k in range(-trad, trad + 1):
        fn = min(max(n + k, 0), n_last)
        if fn != n and fn not in out:
            out.append(fn)
    return out


def trecon_params(clip: VideoNode, **args):
    """tReconCreate's parsing and validation (plugin.cpp:921-976)."""
    _check_args("TRecon", args, TRECON_ARGS)
    fmt = clip.format
    if (fmt.color_family != YUV or (fmt.sample_type == FLOAT and fmt.bits_per_sample != 32)
            or fmt.subsampling_w > 1 or fmt.subsampling_h > 1):
        raise Error("TRecon: planar YUV 444/422/420, 8-16 bit int or 32 bit float")
    if clip.width == 0 or clip.height == 0 or clip.num_frames == 0:
        raise Error("TRecon: constant dimensions and frame count required")
    p = abi.Params()
    abi.lib().lgcr_params_default_trecon(p)
    p.strength = _getf(args, "strength", 0.8)
    p.sigma = _getf(args, "sigma", 0.01)
    p.sratio = _getf(args, "sratio", 0.15)
    p.sdb = _getf(args, "sdb", 3.0)
    p.gsigma = _getf(args, "gsigma", 2.5)
    p.stretch = _getf(args, "stretch", 1.0)
    p.ar_margin = _getf(args, "ar", 0.0)
    p.ms = _getf(args, "ms", 1.0)
    p.trad = int(_get(args, "trad", 1))
    p.tsearch = int(_get(args, "tsearch", 6))
    p.tsad = _getf(args, "tsad", 0.02)
    if (p.strength < 0.0 or p.strength > 1.0 or p.sigma <= 0.0 or p.sratio <= 0.0 or p.sdb <= 0.0
            or p.stretch < 0.0 or p.gsigma <= 0.0 or p.ms < 0.0 or p.ms > 1.0 or p.trad < 0 or p.trad > 8
            or p.tsearch < 0 or p.tsearch > 64 or p.tsad <= 0.0):
        raise Error("TRecon: invalid parameter range (need 0<=strength<=1, "
                    "sigma/sratio/sdb/gsigma/tsad>0, stretch>=0, 0<=trad<=8, 0<=tsearch<=64)")
    p.ridge = 1 if int(_get(args, "ridge", 1)) != 0 else 0
    p.bm = 1 if int(_get(args, "bm", 0)) != 0 else 0
    p.sparse = 1 if int(_get(args, "sparse", 0)) != 0 else 0
    p.algo, p.kernel, p.kp1, p.support, p.radial = 2, abi.KERNELS["lanczos"], 3.0, 3.0, 0
    p.shift_x, p.shift_y, p.loc_set = -0.5, 0.0, 0
    return p


def TRecon(clip: VideoNode, **args) -> VideoNode:
    p = trecon_params(clip, **args)
    fmt_in = clip.format
    fmt_out = replace(fmt_in, subsampling_w=0, subsampling_h=0)
    with _frame_error("TRecon"):
        backend = abi.Clip(_device(), p, _abi_format(fmt_in, clip.width, clip.height), temporal=True)
    n_last = clip.num_frames - 1
    trad = p.trad

    def get(n):
        src = clip.get_frame(n)
        sx, sy = _siting(src.props, fmt_in, -0.5, 0.0, False)
        nbrs = [clip.get_frame(k).planes for k in temporal_frame_numbers(n, n_last, trad)]
        with _frame_error("TRecon"):
            out = backend.trecon_host(src.planes, nbrs, shift_x=sx, shift_y=sy)
        props = dict(src.props)
        props.pop("_ChromaLocation", None)
        return VideoFrame(fmt_out, out, props)

    node = VideoNode(fmt_out, clip.width, clip.height, clip.num_frames, get)
    node._backend = backend
    return node


# ---------------------------------------------------------------------------
# Sharpen (sharpenCreate / sharpenGetFrame)
# ---------------------------------------------------------------------------

SHARPEN_ARGS = ("alpha", "sigma", "sratio", "gspatial", "ar")


def Sharpen(clip: VideoNode, **args) -> VideoNode:
    _check_args("Sharpen", args, SHARPEN_ARGS)
    fmt = clip.format
    if fmt.color_family != YUV or (fmt.sample_type == FLOAT and fmt.bits_per_sample != 32):
        raise Error("LGCR Sharpen: planar YUV only")
    sp = abi.sharpen_params(alpha=_getf(args, "alpha", 0.3), sigma=_getf(args, "sigma", 0.01),
                            sratio=_getf(args, "sratio", 0.15), gspatial=_getf(args, "gspatial", 1.2),
                            ar=_getf(args, "ar", 0.0))
    if sp.sigma <= 0.0 or sp.sratio <= 0.0 or sp.gspatial <= 0.0:
        raise Error("LGCR Sharpen: sigma/sratio/gspatial must be > 0")
    with _frame_error("LGCR Sharpen"):
        backend = abi.SharpenClip(_device(), sp, _abi_format(fmt, clip.width, clip.height))

    def get(n):
        src = clip.get_frame(n)
        with _frame_error("LGCR Sharpen"):
            planes = backend.sharpen_host(*src.planes)
        return VideoFrame(fmt, planes, dict(src.props))

    node = VideoNode(fmt, clip.width, clip.height, clip.num_frames, get)
    node._backend = backend
    return node


# ---------------------------------------------------------------------------
# Downsample (downsampleCreate / downsampleGetFrame, downsample.cpp:725-906)
# ---------------------------------------------------------------------------

DOWNSAMPLE_ARGS = ("quality", "kernel", "strength", "loc")


def _ds_profile_props(props, st, w, h, quality, kernel, strength, loc):
    """writeDownsampleProfile (downsample.cpp:31-49) with the work counts of
    processFrame (downsample.cpp:767-789) and the backend's stage times."""
    cw, ch = w // 2, h // 2
    pixels = 2 * cw * ch
    if kernel == "binomial":
        # 5 taps at integer phase (shift +-0.5), else 4 (downsample.cpp:52-58)
        xsup = 5 if loc in ("left", "topleft", "bottomleft") else 4
        ysup = 4 if loc in ("left", "center") else 5
    else:
        xsup = ysup = 12
    base_taps = 2 * (h * cw * xsup + ch * cw * ysup)
    work = {13: (pixels, base_taps), 16: (pixels, 0)}
    if strength > 0.0:
        diameter = 2 * (quality + 2) + 1
        candidates = (1, 2, 5)[quality]
        work[14] = (pixels, cw * ch * diameter * diameter * candidates)
        if quality == 2:
            work[15] = (pixels, cw * ch * 16)
    for i, name in enumerate(PROP_STAGE_NAMES):
        px, taps = work.get(i, (0, 0))
        props[f"_LGCR_{name}_us"] = float(st.stage_us[i])
        props[f"_LGCR_{name}_pixels"] = px
        props[f"_LGCR_{name}_taps"] = taps
    props["_LGCR_total_us"] = float(st.total_us)
    props["_LGCR_output_pixels"] = w * h + pixels
    props["_LGCR_taps_visited"] = base_taps


def Downsample(clip: VideoNode, **args) -> VideoNode:
    """core.lgcr.Downsample(clip, quality=0, kernel="spline36", strength=1.0,
    loc="left"): planar YUV 4:4:4 to 4:2:0 with the same sample type; the
    output's Y plane is the source's (plane sharing) and ``_ChromaLocation``
    is set from ``loc``."""
    _check_args("Downsample", args, DOWNSAMPLE_ARGS)
    fmt = clip.format
    integer = fmt.sample_type == INTEGER and 8 <= fmt.bits_per_sample <= 16
    floating = fmt.sample_type == FLOAT and fmt.bits_per_sample == 32
    if fmt.color_family != YUV or fmt.subsampling_w != 0 or fmt.subsampling_h != 0 or not (integer or floating):
        raise Error("Downsample: input must be planar YUV444 (8-16 bit integer or float32)")
    if clip.width <= 0 or clip.height <= 0 or clip.width % 2 or clip.height % 2:
        raise Error("Downsample: constant even input width and height are required")
    quality = int(_get(args, "quality", 0))
    if quality < 0 or quality > 2:
        raise Error("Downsample: quality must be 0, 1, or 2")
    kernel = _get(args, "kernel", "spline36")
    if isinstance(kernel, bytes):
        kernel = kernel.decode()
    if kernel not in abi.DS_KERNELS:
        raise Error('Downsample: kernel must be "spline36", "lanczos3", or "binomial"')
    strength = float(np.float32(_get(args, "strength", 1.0)))
    if not math.isfinite(strength) or strength < 0.0 or strength > 1.0:
        raise Error("Downsample: strength must be in the range 0..1")
    loc = _get(args, "loc", "left")
    if isinstance(loc, bytes):
        loc = loc.decode()
    if loc not in abi.DS_LOCATIONS:
        raise Error("Downsample: loc must be left, center, topleft, top, bottomleft, or bottom")
    dp = abi.downsample_params(quality=quality, kernel=kernel, strength=strength, loc=loc)
    with _frame_error("Downsample"):
        backend = abi.DownsampleClip(_device(), dp, _abi_format(fmt, clip.width, clip.height))
    fmt_out = replace(fmt, subsampling_w=1, subsampling_h=1)

    def get(n):
        src = clip.get_frame(n)
        profile = _profiling()
        with _frame_error("Downsample"):
            res = backend.downsample_host(*src.planes, stats=profile, copy_luma=False)
        out, st = res if profile else (res, None)
        props = dict(src.props)
        props["_ChromaLocation"] = abi.DS_LOCATIONS[loc]
        if profile:
            _ds_profile_props(props, st, clip.width, clip.height, quality, kernel, strength, loc)
        return VideoFrame(fmt_out, list(out), props)

    node = VideoNode(fmt_out, clip.width, clip.height, clip.num_frames, get)
    node._backend = backend
    return node


class _Namespace:
    Recon = staticmethod(Recon)
    TRecon = staticmethod(TRecon)
    Sharpen = staticmethod(Sharpen)
    Downsample = staticmethod(Downsample)


class _Core:
    lgcr = _Namespace()


core = _Core()
