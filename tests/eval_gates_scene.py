"""Scenes of the reference's gate study (test/eval_gates.py:55-128) and its
4:2:0 degradations (eval_gates.py:135-188), generated here for the tests.

156 scenes of 128x128 float planes, each with an evaluation band: 150 edges
(softness x chroma shift x angle x contrast), an isoluminant chroma edge,
three Nyquist-alias luma patterns, ridge line art and a ramp.  Each scene is
degraded to 4:2:0 by one of three separable 2x kernels at left siting (box
over luma columns {2c-1, 2c}, triangle, bicubic b=0 c=0.6).  The arithmetic
follows the study's own float64 expressions term by term: the SHA-256 hashes
in tests/golden/eval_gates.npz, written by the study's functions, pin every
scene and degraded plane bit for bit.
"""
import numpy as np

SIZE = 128
_ROWS, _COLS = np.mgrid[0:SIZE, 0:SIZE]
KINDS = ("box", "tri", "bic")


def _vertical_band(halfwidth):
    return (np.abs(_COLS - 64) <= halfwidth) & (_ROWS > 6) & (_ROWS < SIZE - 6)


def _edge(soft, shift, angle, contrast, luma_step=0.05):
    """A luma step at the frame centre with a co-located (shift 0) or
    displaced chroma step, hard (soft 0) or ramped over `soft` pixels."""
    a = np.deg2rad(angle)
    t = (_COLS - 64.0) * np.cos(a) + (_ROWS - 64.0) * np.sin(a) + 64.0
    ty = np.clip(t - 64.0 + 0.5, 0, 1)
    if soft:
        tc = np.clip((t - (64.0 + shift) + soft / 2) / soft, 0, 1)
    else:
        tc = np.clip(t - (64.0 + shift) + 0.5, 0, 1)
    y = (0.30 + luma_step * ty).astype(np.float32)
    u = (-0.15 + contrast * tc).astype(np.float32)
    v = (0.40 - (contrast + 0.05) * tc).astype(np.float32)
    band = (np.abs(t - 64.0) <= 3) & (t != 64.0) & (_ROWS > 6) & (_ROWS < SIZE - 6)
    return y, u, v, band


def scenes():
    """[(name, Y, U, V, band)] in the study's order."""
    out = []
    for soft in (0, 2, 3, 5, 7):
        for shift in (0, 0.5, 1, 2, 4):
            for angle in (0.0, 22.5, 45.0):
                for contrast in (0.55, 0.20):
                    out.append((f"edge_s{soft}_m{shift}_a{int(angle)}_c{contrast}",
                                *_edge(soft, shift, angle, contrast)))
    step = np.clip(_COLS - 64 + 0.5, 0, 1)
    chroma_u = (-0.15 + 0.55 * step).astype(np.float32)
    chroma_v = (0.40 - 0.60 * step).astype(np.float32)
    out.append(("isoluminant", np.full((SIZE, SIZE), 0.35, np.float32), chroma_u, chroma_v, _vertical_band(3)))
    checker = {
        "nullspace_x": (_COLS % 2) == 0,
        "nullspace_y": (_ROWS % 2) == 0,
        "nullspace_xy": ((_COLS + _ROWS) % 2) == 0,
    }
    for name, even in checker.items():
        alias = 0.10 * np.where(even, 1.0, -1.0)
        out.append((name, (0.30 + 0.05 * step + alias).astype(np.float32), chroma_u, chroma_v,
                    _vertical_band(3)))
    dist = np.abs(_COLS - 64)
    ridge_c = np.clip((3 - dist) / 3, 0, 1)
    out.append(("ridge", np.where(dist <= 1, 0.35, 0.70).astype(np.float32),
                (-0.05 + 0.10 * ridge_c).astype(np.float32), np.full((SIZE, SIZE), 0.10, np.float32),
                (dist <= 5) & (_ROWS > 6) & (_ROWS < SIZE - 6)))
    ramp = np.clip((_COLS - 60) / 7.0, 0, 1)
    out.append(("ramp", (0.30 + 0.05 * ramp).astype(np.float32), (-0.15 + 0.55 * ramp).astype(np.float32),
                (0.40 - 0.60 * ramp).astype(np.float32),
                (_COLS >= 56) & (_COLS < 72) & (_ROWS > 6) & (_ROWS < SIZE - 6)))
    return out


def _keys_weight(d, b=0.0, c=0.6):
    """Mitchell-Netravali cubic (b, c) at |d| (in output-sample units)."""
    if d < 1:
        return ((12 - 9 * b - 6 * c) * d**3 + (-18 + 12 * b + 6 * c) * d**2 + (6 - 2 * b)) / 6
    if d < 2:
        return ((-b - 6 * c) * d**3 + (6 * b + 30 * c) * d**2 + (-12 * b - 48 * c) * d + (8 * b + 24 * c)) / 6
    return 0.0


def _taps(n, kind, siting):
    """Per output sample of a 2x decimation: [(clamped source index, weight)]
    in increasing source order, weights normalised to sum 1."""
    rows = []
    for c in range(n // 2):
        centre = (c + 0.5) * 2 - 0.5 + siting
        acc = {}
        if kind == "box":
            for i in range(int(np.ceil(centre - 1)), int(np.ceil(centre + 1))):
                acc[i] = acc.get(i, 0.0) + 0.5
        else:
            reach = 2.0 if kind == "tri" else 4.0
            for i in range(int(np.floor(centre - reach)), int(np.ceil(centre + reach)) + 1):
                d = abs(i - centre) / 2.0
                w = max(0.0, 1.0 - d) if kind == "tri" else _keys_weight(d)
                if w:
                    acc[i] = acc.get(i, 0.0) + w
        total = sum(acc.values())
        rows.append([(min(max(i, 0), n - 1), w / total) for i, w in acc.items()])
    return rows


def degrade(plane, kind):
    """Float64 plane -> 4:2:0 chroma grid (horizontal pass, then vertical),
    left siting: horizontal centre at luma 2c, vertical centre between rows."""
    h, w = plane.shape
    mid = np.zeros((h, w // 2))
    for cx, taps in enumerate(_taps(w, kind, -0.5)):
        for i, wt in taps:
            mid[:, cx] += plane[:, i] * wt
    out = np.zeros((h // 2, w // 2))
    for cy, taps in enumerate(_taps(h, kind, 0.0)):
        for j, wt in taps:
            out[cy, :] += mid[j, :] * wt
    return out


def observations():
    """Yield (scene index, name, kind, Y, U, V, band, U420, V420): the planes
    the plugin sees (float32) and the float64 source chroma for the error."""
    for si, (name, y, u, v, band) in enumerate(scenes()):
        u64, v64 = u.astype(np.float64), v.astype(np.float64)
        for kind in KINDS:
            yield si, name, kind, y, u, v, band, degrade(u64, kind), degrade(v64, kind)
