"""Shared helpers for the GPU-vs-oracle parity tests."""
import numpy as np

import oracle
from lgcr_amd import abi, synth

FLOAT_TOL = 1e-5  # absolute, float outputs
INT_TOL = 1       # code values, integer outputs


def run_both(y, u, v, ss_w=1, ss_h=1, bits=None, device=None, **params):
    """Run one frame through the oracle and through the HIP backend (host
    planes via the C ABI).  Returns (oracle_out, oracle_debug, gpu_out, clip)."""
    h, w = y.shape
    ow = params.get("width", w)
    oh = params.get("height", h)
    op = {k: v for k, v in params.items() if k not in ("width", "height")}
    p_or = oracle.make_params(ow, oh, **op)
    o_out = oracle.recon_frame(p_or, y, u, v, bits=bits, ss_w=ss_w, ss_h=ss_h)
    o_dbg = oracle.last_debug()
    fmt = abi.make_format(y.dtype, w, h, ss_w, ss_h, bits)
    gp = abi.default_params(**params)
    gp.shift_x = p_or.shift_x if ss_w else 0.0
    gp.shift_y = p_or.shift_y if ss_h else 0.0
    clip = abi.Clip(device, gp, fmt)
    g_out = clip.recon_host(y, u, v)
    return o_out, o_dbg, g_out, clip


def compare_planes(a, b, is_float):
    """(mismatch count, max abs diff) with the north-star tolerance."""
    if is_float:
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        return int((d > FLOAT_TOL).sum()), float(d.max()) if d.size else 0.0
    d = np.abs(a.astype(np.int64) - b.astype(np.int64))
    return int((d > INT_TOL).sum()), float(d.max()) if d.size else 0.0


def frame(width, height, ss_w=1, ss_h=1, bits=8, seed=0, frame=0, **kw):
    return synth.clip_frame(width, height, ss_w=ss_w, ss_h=ss_h, bits=bits, seed=seed, frame=frame, **kw)
