"""The reference's quality battery (test/battery.py), run on this backend.

Every case builds a 128x128 4:4:4 float ground truth, degrades it to 4:2:0
the way the battery does (VapourSynth's resize.Bilinear, test/battery.py:
46-47), reconstructs it and measures the chroma MAE in the case's band
(:50-53, :56-69).  The table, the temporal cases and the integer round trip
print in the layout of the reference's test/results/latest.md, and
``--check`` applies the battery's quality gates (:260-290).

The degradation restates the published algorithm of zimg's resizer, which
VapourSynth's resize.Bilinear runs (zimg is not part of the reference):
per output sample i and axis, the position on the source grid is
(i + 0.5) / scale + shift; the filter support grows by 1 / scale when
downscaling; 2 * ceil(support) taps start at floor(pos + support - taps) +
0.5; each tap weighs filter((x - pos) * min(scale, 1)) divided by the sum of
the taps; positions before the first sample mirror to -x and past the last
to min(2n - x, n - 0.5); float samples accumulate tap by tap in float.
4:4:4 -> 4:2:0 with the default left chroma siting moves the horizontal
position by -0.5 source samples and the vertical by 0 (bilinear at 2:1:
taps 1/4, 1/2, 1/4 around luma column 2c; 1/8, 3/8, 3/8, 1/8 over rows
2r-1 .. 2r+2).  The upscale case's reference picture uses zimg's Spline36
the same way.

Run: python -m lgcr_amd.battery [--all] [--check]
"""
from __future__ import annotations

import math
import sys

import numpy as np

W = H = 128
_YY, _XX = np.mgrid[0:H, 0:W]


# ---------------------------------------------------------------------------
# resampling (zimg's resize filter, restated)
# ---------------------------------------------------------------------------

def _triangle(x):
    x = abs(x)
    return 1.0 - x if x < 1.0 else 0.0


def _spline36(x):
    x = abs(x)
    if x < 1.0:
        return ((13.0 / 11.0 * x - 453.0 / 209.0) * x - 3.0 / 209.0) * x + 1.0
    if x < 2.0:
        x -= 1.0
        return ((-6.0 / 11.0 * x + 270.0 / 209.0) * x - 156.0 / 209.0) * x
    if x < 3.0:
        x -= 2.0
        return ((1.0 / 11.0 * x - 45.0 / 209.0) * x + 26.0 / 209.0) * x
    return 0.0


def filter_taps(kernel, support, src, dst, shift):
    """Per output sample: [(source index, weight)] in ascending index order."""
    scale = dst / src
    step = min(scale, 1.0)
    sup = support / step
    size = max(int(math.ceil(sup)) * 2, 1)
    rows = []
    for i in range(dst):
        pos = (i + 0.5) / scale + shift
        begin = math.floor(pos + sup - size) + 0.5
        total = sum(kernel((begin + j - pos) * step) for j in range(size))
        acc = {}
        for j in range(size):
            x = begin + j
            if x < 0.0:
                real = -x
            elif x >= src:
                real = min(2.0 * src - x, src - 0.5)
            else:
                real = x
            k = int(math.floor(real))
            acc[k] = acc.get(k, 0.0) + kernel((x - pos) * step) / total
        rows.append(sorted(acc.items()))
    return rows


def _pass(plane, taps, axis):
    """One resampling pass over the columns (axis 1) or rows (axis 0) of a
    float32 plane, accumulating in float32 tap by tap."""
    src = plane if axis == 1 else plane.T
    out = np.empty((src.shape[0], len(taps)), np.float32)
    for i, row in enumerate(taps):
        acc = np.zeros(src.shape[0], np.float32)
        for k, w in row:
            acc = (acc + np.float32(w) * src[:, k]).astype(np.float32)
        out[:, i] = acc
    return out if axis == 1 else np.ascontiguousarray(out.T)


def degrade420(y, u, v):
    """4:4:4 float -> 4:2:0 with left chroma siting (resize.Bilinear)."""
    th = filter_taps(_triangle, 1.0, y.shape[1], y.shape[1] // 2, -0.5)
    tv = filter_taps(_triangle, 1.0, y.shape[0], y.shape[0] // 2, 0.0)
    return y, _pass(_pass(u, th, 1), tv, 0), _pass(_pass(v, th, 1), tv, 0)


def spline36_resize(plane, ow, oh):
    """A float plane resized with Spline36 (resize.Spline36, same siting)."""
    th = filter_taps(_spline36, 3.0, plane.shape[1], ow, 0.0)
    tv = filter_taps(_spline36, 3.0, plane.shape[0], oh, 0.0)
    return _pass(_pass(plane, th, 1), tv, 0)


# ---------------------------------------------------------------------------
# cases (test/battery.py:86-205)
# ---------------------------------------------------------------------------

def _edge(xedg, soft=0, cshift=0):
    """Small luma step at xedg, chroma step at xedg + cshift, hard or ramped
    over `soft` pixels."""
    ty = np.clip(_XX - xedg + 0.5, 0, 1)
    if soft:
        tc = np.clip((_XX - (xedg + cshift) + soft / 2) / soft, 0, 1)
    else:
        tc = np.clip(_XX - (xedg + cshift) + 0.5, 0, 1)
    return ((0.30 + 0.05 * ty).astype(np.float32), (-0.15 + 0.55 * tc).astype(np.float32),
            (0.40 - 0.60 * tc).astype(np.float32))


def _rows_inside(mask):
    return mask & (_YY > 4) & (_YY < H - 4)


def cases():
    """{name: (Y, U, V, band, scale)} in the battery's order."""
    out = {}
    band_v = _rows_inside((np.abs(_XX - 64) <= 3) & (_XX != 64))
    out["hard_v"] = (*_edge(64), band_v, 1.0)
    t = np.clip(_XX - _YY + 0.5, 0, 1)
    out["hard_d45"] = ((0.30 + 0.05 * t).astype(np.float32), (-0.15 + 0.55 * t).astype(np.float32),
                       (0.40 - 0.60 * t).astype(np.float32),
                       _rows_inside((np.abs(_XX - _YY) <= 2) & (_XX != _YY)), 1.0)
    _, u, v = _edge(64)
    out["isoluminant"] = (np.full((H, W), 0.35, np.float32), u, v, band_v, 1.0)
    out["misalign4"] = (*_edge(64, cshift=4), _rows_inside(np.abs(_XX - 66) <= 5), 1.0)
    out["misalign1"] = (*_edge(64, cshift=1), _rows_inside(np.abs(_XX - 65) <= 4), 1.0)
    step = np.clip(_XX - 64 + 0.5, 0, 1)
    for name, even in (("nullspace_x", (_XX % 2) == 0), ("nullspace_y", (_YY % 2) == 0),
                       ("nullspace_xy", ((_XX + _YY) % 2) == 0)):
        alias = 0.10 * np.where(even, 1.0, -1.0)
        out[name] = ((0.30 + 0.05 * step + alias).astype(np.float32), (-0.15 + 0.55 * step).astype(np.float32),
                     (0.40 - 0.60 * step).astype(np.float32), band_v, 1.0)
    out["hardL_softC"] = (*_edge(64, soft=7), band_v, 1.0)
    dist = np.abs(_XX - 64)
    tc = np.clip((3 - dist) / 3, 0, 1)
    out["ridge_line"] = (np.where(dist <= 1, 0.35, 0.70).astype(np.float32), (-0.05 + 0.10 * tc).astype(np.float32),
                         np.full((H, W), 0.10, np.float32), _rows_inside(dist <= 5), 1.0)
    t = np.clip((_XX - 60) / 7.0, 0, 1)
    out["ramp"] = ((0.30 + 0.05 * t).astype(np.float32), (-0.15 + 0.55 * t).astype(np.float32),
                   (0.40 - 0.60 * t).astype(np.float32), _rows_inside((_XX >= 58) & (_XX < 70)), 1.0)
    out["texture"] = (np.full((H, W), 0.5, np.float32),
                      (0.15 * np.sin(_XX * 0.9) * np.sin(_YY * 0.7)).astype(np.float32),
                      np.zeros((H, W), np.float32), (_XX >= 8) & (_XX < 120) & (_YY >= 8) & (_YY < 120), 1.0)
    rng = np.random.default_rng(7)
    y, u, v = _edge(64)
    y = np.clip(y + rng.normal(0, 0.008, (H, W)), 0, 1).astype(np.float32)
    u = np.clip(u + rng.normal(0, 0.004, (H, W)), -0.5, 0.5).astype(np.float32)
    out["noise"] = (y, u, v, band_v, 1.0)
    out["upscale2x"] = (*_edge(64), band_v, 2.0)
    return out


def _band_mae(planes, gu, gv, mask):
    a = np.abs(planes[1].astype(np.float64) - gu.astype(np.float64))[mask].mean()
    b = np.abs(planes[2].astype(np.float64) - gv.astype(np.float64))[mask].mean()
    return float((a + b) / 2)


# ---------------------------------------------------------------------------
# runs.  recon(kwargs, (Y, U, V), out_w, out_h) and trecon(kwargs, center,
# neighbours) -> output planes; kwargs are Recon / TRecon arguments.
# ---------------------------------------------------------------------------

def algos(all_algos=False):
    a = {"plain": dict(strength=0.0), "algo2": dict(strength=0.8, algo=2)}
    if all_algos:
        for k in (4, 6):
            a[f"algo{k}"] = dict(strength=0.8, algo=k)
    return a


def run_cases(recon, tags):
    """{case: {tag: band MAE}} for the table (test/battery.py:56-69)."""
    res = {}
    for name, (y, u, v, mask, scale) in cases().items():
        src = degrade420(y, u, v)
        ow, oh = int(W * scale), int(H * scale)
        gu, gv = u, v
        if scale != 1.0:
            gu, gv = spline36_resize(u, ow, oh), spline36_resize(v, ow, oh)
            mask = np.kron(mask, np.ones((2, 2))) > 0
        res[name] = {tag: _band_mae(recon(dict(kernel="jinc", taps=3, **kw), src, ow, oh), gu, gv, mask)
                     for tag, kw in tags.items()}
    return res


def run_temporal(recon, tags):
    """Aligned edge-band error variance of a one-pixel move (:208-238)."""
    frames = []
    for k in range(2):
        t = np.clip(_XX - (64 + k) + 0.5, 0, 1)
        frames.append(((0.30 + 0.05 * t).astype(np.float32), (-0.15 + 0.55 * t).astype(np.float32),
                       (0.40 - 0.60 * t).astype(np.float32)))
    band0 = _rows_inside(np.abs(_XX - 64) <= 4) & (_XX < W - 1)
    out = {}
    for tag, kw in tags.items():
        rec = [recon(dict(kernel="jinc", taps=3, **kw), degrade420(*f), W, H) for f in frames]
        var = []
        for p in (1, 2):
            e0 = rec[0][p].astype(np.float64) - frames[0][p]
            e1 = rec[1][p].astype(np.float64) - frames[1][p]
            var.append(np.var(np.stack((e0, np.roll(e1, -1, axis=1))), axis=0)[band0].mean())
        out[tag] = float(np.mean(var))
    return out


def _temporal_edge(xedg, noise=0.0, seed=0):
    t = np.clip(_XX - xedg + 0.5, 0, 1)
    y = (0.30 + 0.05 * t).astype(np.float32)
    u = (-0.15 + 0.55 * t).astype(np.float32)
    v = (0.40 - 0.60 * t).astype(np.float32)
    if noise:
        rng = np.random.default_rng(seed)
        y = np.clip(y + rng.normal(0, noise, (H, W)), 0, 1).astype(np.float32)
        u = np.clip(u + rng.normal(0, noise, (H, W)), -0.5, 0.5).astype(np.float32)
    return y, u, v


def run_trecon_cases(recon, trecon):
    """Odd- and even-pel motion, static noise, single-frame isolation and flat
    luma (:241-318): frame 1 of three, TRecon with radius 1."""
    band = _rows_inside(np.abs(_XX - 64) <= 4)
    out = {}
    for label, frames in (("t_move1", [_temporal_edge(x) for x in (63, 64, 65)]),
                          ("t_move2", [_temporal_edge(x) for x in (62, 64, 66)]),
                          ("t_move3", [_temporal_edge(x) for x in (61, 64, 67)]),
                          ("t_static_noise", [_temporal_edge(64, 0.01, s) for s in (1, 2, 3)])):
        src = [degrade420(*f) for f in frames]
        gt = frames[1]
        outs = {"plain": recon(dict(strength=0.0), src[1], W, H),
                "algo2": recon(dict(strength=0.8, algo=2), src[1], W, H),
                "trecon": trecon(dict(strength=0.8, trad=1), src[1], [src[0], src[2]])}
        out[label] = {tag: float(np.mean([np.abs(o[p].astype(np.float64) - gt[p].astype(np.float64))[band].mean()
                                          for p in (1, 2)])) for tag, o in outs.items()}
    src = degrade420(*_temporal_edge(64))
    fr = recon(dict(strength=0.8, algo=2), src, W, H)
    ft = trecon(dict(strength=0.8, trad=1), src, [])
    out["t_single"] = {"isolation_maxdiff": float(max(np.abs(fr[p].astype(np.float64) - ft[p].astype(np.float64))
                                                      .max() for p in (1, 2)))}
    y = np.full((H, W), 0.5, np.float32)
    u = (0.15 * np.sin(_XX * 0.9) * np.sin(_YY * 0.7)).astype(np.float32)
    v = (0.10 * np.cos(_XX * 0.5)).astype(np.float32)
    src = degrade420(y, u, v)
    outs = {"algo2": recon(dict(strength=0.8, algo=2, sparse=0), src, W, H),
            "trecon": trecon(dict(strength=0.8, trad=1), src, [src, src])}
    out["t_flat_amb"] = {tag: float(np.mean([np.abs(o[p].astype(np.float64) - g.astype(np.float64)).mean()
                                             for p, g in ((1, u), (2, v))])) for tag, o in outs.items()}
    return out


def run_int8(recon, trecon):
    """Constant YUV420P8 (64, 128, 128) through Recon and TRecon (:321-338)."""
    y = np.full((64, 64), 64, np.uint8)
    u = np.full((32, 32), 128, np.uint8)
    v = np.full((32, 32), 128, np.uint8)
    res = {"recon8": [int(p[32, 32]) for p in recon({}, (y, u, v), 64, 64)],
           "trecon8": [int(p[32, 32]) for p in trecon({}, (y, u, v), [(y, u, v), (y, u, v)])]}
    ok = all(abs(a - b) <= 1 for vals in res.values() for a, b in zip(vals, (64, 128, 128)))
    return res, ok


def report(case_res, temporal, trc, int8):
    """The text of test/results/latest.md (:341-381)."""
    tags = list(next(iter(case_res.values())).keys())
    header = f"{'case':<13}" + "".join(f"{t:>10}" for t in tags)
    lines = [header, "-" * len(header)]
    for name, row in case_res.items():
        lines.append(f"{name:<13}" + "".join(f"{row[t]:>10.5f}" for t in tags))
    lines.append(f"{'temporal':<13}" + "".join(f"{temporal[t]:>10.5f}" for t in tags)
                 + "   (aligned edge-band error variance; lower=better)")
    lines.append("")
    lines.append("temporal cases (plain / algo2 / trecon):")
    for label, row in trc.items():
        if label == "t_single":
            lines.append(f"{label:<13} isolation maxdiff = {row['isolation_maxdiff']:.6f}"
                         "  (must be ~0: single frame has no temporal information)")
        elif label == "t_flat_amb":
            lines.append(f"{label:<13}" + "".join(f"{row[t]:>10.5f}" for t in ("algo2", "trecon"))
                         + "   (identical frames, unobservable motion: trecon ~= algo2 required)")
        else:
            lines.append(f"{label:<13}" + "".join(f"{row[t]:>10.5f}" for t in ("plain", "algo2", "trecon")))
    res8, ok8 = int8
    lines.append("")
    lines.append(f"int8 constant (64,128,128): {res8}  {'OK' if ok8 else 'FAIL'}")
    return "\n".join(lines)


def check_gates(case_res, trc, int8_ok):
    """The battery's quality gates (test/battery.py:260-290); raises
    AssertionError on the first that fails."""
    assert int8_ok, "integer Recon/TRecon round-trip failed"
    assert trc["t_single"]["isolation_maxdiff"] <= 1e-6, trc["t_single"]
    assert abs(trc["t_move2"]["trecon"] - trc["t_move2"]["algo2"]) <= 1e-6, trc["t_move2"]
    assert trc["t_flat_amb"]["trecon"] <= trc["t_flat_amb"]["algo2"] + 1e-6, trc["t_flat_amb"]
    assert trc["t_static_noise"]["trecon"] < trc["t_static_noise"]["algo2"], trc["t_static_noise"]
    for name in ("hard_v", "hard_d45", "upscale2x"):
        assert case_res[name]["algo2"] <= case_res[name]["plain"] + 1e-6, name
    for name in ("isoluminant", "misalign4", "texture"):
        assert abs(case_res[name]["algo2"] - case_res[name]["plain"]) <= 1e-6, name
    for name in ("hardL_softC", "ridge_line", "ramp", "noise"):
        assert case_res[name]["algo2"] <= case_res[name]["plain"] + 0.001, name


def run_all(recon, trecon, all_algos=True):
    tags = algos(all_algos)
    case_res = run_cases(recon, tags)
    temporal = run_temporal(recon, tags)
    trc = run_trecon_cases(recon, trecon)
    int8 = run_int8(recon, trecon)
    return case_res, temporal, trc, int8


# ---------------------------------------------------------------------------
# the HIP backend as the battery's Recon / TRecon
# ---------------------------------------------------------------------------

def hip_runners(device):
    """(recon, trecon, close) backed by the C ABI; clips are cached per
    argument set and format."""
    from . import abi
    clips = {}

    def fmt_of(y):
        return abi.make_format(y.dtype, y.shape[1], y.shape[0])

    def clip(kind, kw, planes, ow, oh):
        y = planes[0]
        key = (kind, tuple(sorted(kw.items())), y.dtype.str, y.shape, ow, oh)
        if key not in clips:
            if kind == "trecon":
                clips[key] = abi.Clip(device, abi.trecon_params(**kw), fmt_of(y), temporal=True)
            else:
                p = dict(kw)
                if (ow, oh) != (y.shape[1], y.shape[0]):
                    p.update(width=ow, height=oh)
                clips[key] = abi.Clip(device, abi.default_params(**p), fmt_of(y))
        return clips[key]

    def recon(kw, planes, ow, oh):
        return clip("recon", kw, planes, ow, oh).recon_host(*planes)

    def trecon(kw, center, nbrs):
        return clip("trecon", kw, center, center[0].shape[1], center[0].shape[0]).trecon_host(center, list(nbrs))

    def close():
        for c in clips.values():
            c.close()

    return recon, trecon, close


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    from . import abi
    dev = abi.Device(0)
    recon, trecon, close = hip_runners(dev)
    try:
        case_res, temporal, trc, int8 = run_all(recon, trecon, all_algos="--all" in argv)
        print(report(case_res, temporal, trc, int8))
        if "--check" in argv:
            check_gates(case_res, trc, int8[1])
            print("battery gates: OK")
    finally:
        close()
        dev.close()


if __name__ == "__main__":
    main()
