"""Deterministic synthetic clips (polygons, outlines, noise) for tests and
benchmarks; a ctypes view of ``liblgcr_synth.so`` (csrc/synth.c)."""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import paths


class Spec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("width", C.c_int), ("height", C.c_int),
        ("ss_w", C.c_int), ("ss_h", C.c_int), ("bits", C.c_int), ("is_float", C.c_int),
        ("polygons", C.c_int), ("motion", C.c_int), ("scene_per_frame", C.c_int),
        ("luma_noise", C.c_double), ("chroma_noise", C.c_double),
    ]


_lib = None


def _load():
    global _lib
    if _lib is None:
        path = paths.SYNTH_LIB
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.lgcr_synth_frame.argtypes = [C.POINTER(Spec), C.c_int, C.c_void_p * 3, C.c_ssize_t * 3]
        L.lgcr_synth_frame.restype = C.c_int
        _lib = L
    return _lib


def clip_frame(width, height, ss_w=1, ss_h=1, bits=8, frame=0, seed=0, polygons=40,
               motion=False, scene_per_frame=False, luma_noise=None, chroma_noise=0.0):
    """One frame as (Y, U, V) numpy planes.

    Integer formats are uint8 (bits == 8) or uint16; bits == 32 is float.
    luma_noise defaults to 1 LSB for integers and 0.002 for float.
    """
    is_float = bits == 32
    if luma_noise is None:
        luma_noise = 0.002 if is_float else 1.0
    dt = np.float32 if is_float else (np.uint8 if bits <= 8 else np.uint16)
    cw, ch = width >> ss_w, height >> ss_h
    y = np.empty((height, width), dt)
    u = np.empty((ch, cw), dt)
    v = np.empty((ch, cw), dt)
    spec = Spec(seed, width, height, ss_w, ss_h, bits, int(is_float), polygons, int(motion),
                int(scene_per_frame), float(luma_noise), float(chroma_noise))
    planes = (C.c_void_p * 3)(y.ctypes.data, u.ctypes.data, v.ctypes.data)
    strides = (C.c_ssize_t * 3)(y.strides[0], u.strides[0], v.strides[0])
    if _load().lgcr_synth_frame(C.byref(spec), frame, planes, strides) != 0:
        raise ValueError("invalid synthetic clip spec")
    return y, u, v


def clip(frames, width, height, threads=8, **kw):
    """`frames` frames of one clip, generated in parallel (ctypes drops the GIL)."""
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda n: clip_frame(width, height, frame=n, **kw), range(frames)))
