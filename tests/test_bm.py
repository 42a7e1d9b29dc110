"""Collaborative chroma filter (bm=True, bm.cpp:250-442): oracle checks.

The oracle's lifting-form Haar shrinkage is pinned against an independent
float64 model written with the textbook orthonormal Haar matrices, a stable
sort for the match list and float64 window-weighted accumulation; the properties of
the reference's own test/test_bm.py (luma untouched, constant chroma
preserved, noise variance reduced, clean co-edge not regressed, resize and
integer paths finite and constant-preserving) are checked on the oracle's
whole-frame Recon and TRecon.
"""
import numpy as np
import pytest

import oracle

S2 = np.sqrt(0.5)


def haar_matrix(n):
    """Orthonormal Haar analysis matrix, rows ordered coarse to fine."""
    if n == 1:
        return np.ones((1, 1))
    h = haar_matrix(n // 2)
    top = np.kron(h, [1.0, 1.0]) * S2
    bottom = np.kron(np.eye(n // 2), [1.0, -1.0]) * S2
    return np.vstack([top, bottom])


H8 = haar_matrix(8)
WIN = np.array([0.25, 0.5, 0.75, 1.0, 1.0, 0.75, 0.5, 0.25])


def starts(extent):
    s = list(range(0, extent - 8 + 1, 8))
    if s[-1] != extent - 8:
        s.append(extent - 8)
    return s


def bm_model(g, u, v, sigma):
    """float64 model of the collaborative filter (same guide size)."""
    h, w = u.shape
    g, u, v = (a.astype(np.float64) for a in (g, u, v))
    su, sv = np.zeros_like(u), np.zeros_like(v)
    sw = np.zeros_like(u)
    sig = min(max(0.6 * sigma, 0.002), 0.012)
    thr = 2.7 * sig
    limit = max(4e-5, 4.0 * sigma * sigma)
    for ay in starts(h):
        for ax in starts(w):
            cands = []
            for dy in range(-8, 9, 4):
                for dx in range(-8, 9, 4):
                    b = (min(max(ax + dx, 0), w - 8), min(max(ay + dy, 0), h - 8))
                    if b not in cands:
                        cands.append(b)

            def cost(b):
                bx, by = b
                if (bx, by) == (ax, ay):
                    return 0.0
                sl = (slice(0, 8, 2), slice(0, 8, 2))
                pa = lambda p: p[ay:ay + 8, ax:ax + 8][sl]
                pb = lambda p: p[by:by + 8, bx:bx + 8][sl]
                d = (pa(g) - pb(g)) ** 2 + 0.25 * ((pa(u) - pb(u)) ** 2 + (pa(v) - pb(v)) ** 2)
                return d.mean()

            costs = [cost(b) for b in cands]
            order = sorted(range(len(cands)), key=lambda i: (costs[i], i))[:4]
            elig = sum(1 for i in order if costs[i] <= limit)
            gc = 4 if elig >= 4 else (2 if elig >= 2 else 1)
            if gc < 2:
                continue
            pos = [cands[i] for i in order[:gc]]
            hg = haar_matrix(gc)
            grp, lo, hi = [], [], []
            retained = 0
            for p in (u, v):
                stack = np.stack([p[by:by + 8, bx:bx + 8] for bx, by in pos])
                lo.append(stack.min(0))
                hi.append(stack.max(0))
                coef = np.einsum("ij,kjl,ml->kim", H8, stack, H8)       # H P H^T per member
                gcoef = np.einsum("ak,kim->aim", hg, coef)              # across the group
                keep = np.abs(gcoef) >= thr
                keep[0, 0, 0] = True
                retained += int(keep.sum())
                gcoef = np.where(keep, gcoef, 0.0)
                coef = np.einsum("ak,aim->kim", hg, gcoef)
                grp.append(np.einsum("ij,kim,ml->kjl", H8, coef, H8))
            wgt = np.outer(WIN, WIN) / (1.0 + retained)
            for k, (bx, by) in enumerate(pos):
                for c, acc in enumerate((su, sv)):
                    acc[by:by + 8, bx:bx + 8] += wgt * np.clip(grp[c][k], lo[c], hi[c])
                sw[by:by + 8, bx:bx + 8] += wgt
    ou, ov = u.copy(), v.copy()
    m = sw > 0
    ou[m] = u[m] + 0.65 * (su[m] / sw[m] - u[m])
    ov[m] = v[m] + 0.65 * (sv[m] / sw[m] - v[m])
    return ou, ov


def noisy_planes(w, h, seed, sigma=0.006, structure=True):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    g = np.full((h, w), 0.45, np.float32)
    u = np.full((h, w), 0.08, np.float32)
    v = np.full((h, w), -0.06, np.float32)
    if structure:
        disk = (xx - w * 0.4) ** 2 + (yy - h * 0.5) ** 2 < (0.3 * min(w, h)) ** 2
        g[disk] = 0.7
        u[disk] = -0.15
        v[disk] = 0.2
    u = (u + rng.normal(0, sigma, (h, w))).astype(np.float32)
    v = (v + rng.normal(0, sigma, (h, w))).astype(np.float32)
    g = (g + rng.normal(0, 0.002, (h, w))).astype(np.float32)
    return g, u, v


@pytest.mark.parametrize("w,h,seed", [(64, 64, 1), (44, 36, 2), (8, 8, 3), (17, 9, 4)])
def test_oracle_matches_matrix_model(w, h, seed):
    g, u, v = noisy_planes(w, h, seed)
    for sigma in (0.01, 0.003):
        ou, ov = oracle.collaborative(g, u, v, sigma)
        mu, mv = bm_model(g, u, v, sigma)
        assert np.abs(ou - mu).max() < 2e-6
        assert np.abs(ov - mv).max() < 2e-6
        if w == 8 and h == 8:
            # a single candidate patch: no group, no change
            assert np.array_equal(ou, u)
        if w == 64:
            assert not np.array_equal(ou, u)


def test_haar_model_is_orthonormal():
    for n in (2, 4, 8):
        hm = haar_matrix(n)
        assert np.allclose(hm @ hm.T, np.eye(n), atol=1e-15)
    # the DC row of the 8-point transform is the scaled mean
    assert np.allclose(H8[0], np.full(8, 8 ** -0.5))


def test_small_planes_untouched():
    g, u, v = noisy_planes(7, 30, 5)
    ou, ov = oracle.collaborative(g, u, v, 0.01)
    assert np.array_equal(ou, u) and np.array_equal(ov, v)


def test_resized_guide_matches_explicit_bilinear():
    g, u, v = noisy_planes(48, 32, 6)
    big_u = np.repeat(np.repeat(u, 2, 0), 2, 1)[:56, :80].copy()
    big_v = np.repeat(np.repeat(v, 2, 0), 2, 1)[:56, :80].copy()
    # bilinear resize of the guide, lgcr.h:467-475 with clamped taps
    h, w = big_u.shape
    sx, sy = g.shape[1] / w, g.shape[0] / h
    gr = np.zeros((h, w), np.float32)
    for y in range(h):
        fy0 = (y + 0.5) * sy - 0.5
        y0 = int(np.floor(fy0))
        fy = np.float32(fy0 - y0)
        for x in range(w):
            fx0 = (x + 0.5) * sx - 0.5
            x0 = int(np.floor(fx0))
            fx = np.float32(fx0 - x0)
            at = lambda xx, yy: g[min(max(yy, 0), g.shape[0] - 1), min(max(xx, 0), g.shape[1] - 1)]
            a, b, c, d = at(x0, y0), at(x0 + 1, y0), at(x0, y0 + 1), at(x0 + 1, y0 + 1)
            gr[y, x] = a + (b - a) * fx + (c - a) * fy + (a - b - c + d) * fx * fy
    ou1, ov1 = oracle.collaborative(g, big_u, big_v, 0.01)
    ou2, ov2 = oracle.collaborative(gr, big_u, big_v, 0.01)
    assert np.array_equal(ou1, ou2) and np.array_equal(ov1, ov2)


# ---- the reference's own test/test_bm.py properties on whole frames --------

def recon(y, u, v, ss_w=0, ss_h=0, width=None, height=None, **kw):
    h, w = y.shape
    p = oracle.make_params(width or w, height or h, **kw)
    return oracle.recon_frame(p, y, u, v, ss_w=ss_w, ss_h=ss_h)


def noisy_clip_frame(w=64, h=64):
    rng = np.random.default_rng(107)
    nu = rng.normal(0.0, 0.006, (h, w)).astype(np.float32)
    nv = rng.normal(0.0, 0.006, (h, w)).astype(np.float32)
    y = np.full((h, w), 0.45, np.float32)
    return y, (0.08 + nu).astype(np.float32), (-0.06 + nv).astype(np.float32)


def chroma_var(out):
    return float(np.mean([np.var(out[p].astype(np.float64)) for p in (1, 2)]))


def test_recon_default_off_and_denoising():
    y, u, v = noisy_clip_frame()
    implicit = recon(y, u, v, strength=0.0)
    disabled = recon(y, u, v, strength=0.0, bm=0)
    enabled = recon(y, u, v, strength=0.0, bm=1)
    for p in range(3):
        assert np.array_equal(implicit[p], disabled[p])
        assert np.isfinite(enabled[p]).all()
    assert np.array_equal(implicit[0], enabled[0])
    assert chroma_var(enabled) < 0.85 * chroma_var(disabled)


def test_constant_preservation():
    y = np.full((32, 40), 0.4, np.float32)
    u = np.full((32, 40), 0.1, np.float32)
    v = np.full((32, 40), -0.1, np.float32)
    off = recon(y, u, v, strength=0.0, bm=0)
    on = recon(y, u, v, strength=0.0, bm=1)
    assert max(float(np.abs(off[p] - on[p]).max()) for p in (1, 2)) <= 1e-6


def test_trecon_switch():
    y, u, v = noisy_clip_frame()
    nb = [(y, u, v), (y, u, v)]
    off = oracle.trecon_frame(oracle.trecon_params(strength=0.0, trad=1), (y, u, v), nb, ss_w=0, ss_h=0)
    on = oracle.trecon_frame(oracle.trecon_params(strength=0.0, trad=1, bm=1), (y, u, v), nb, ss_w=0, ss_h=0)
    for p in range(3):
        assert np.isfinite(on[p]).all()
    assert np.array_equal(off[0], on[0])
    assert chroma_var(on) < 0.85 * chroma_var(off)


def test_edge_preservation():
    w = h = 64
    yy, xx = np.mgrid[0:h, 0:w]
    side = xx >= w // 2
    exp = [np.where(side, a, b).astype(np.float32) for a, b in ((0.58, 0.24), (0.22, -0.18), (-0.20, 0.16))]
    # 2x2 box down to 4:2:0 (the edge sits on an even column, so each
    # chroma sample is one side's value)
    cu = exp[1].reshape(h // 2, 2, w // 2, 2).mean((1, 3)).astype(np.float32)
    cv = exp[2].reshape(h // 2, 2, w // 2, 2).mean((1, 3)).astype(np.float32)
    off = recon(exp[0], cu, cv, ss_w=1, ss_h=1, algo=2, sparse=0, bm=0)
    on = recon(exp[0], cu, cv, ss_w=1, ss_h=1, algo=2, sparse=0, bm=1)
    band = np.abs(xx - w // 2) <= 4

    def mae(o):
        return float(np.mean([np.abs(o[p].astype(np.float64) - exp[p])[band].mean() for p in (1, 2)]))

    assert mae(on) <= mae(off) + 1e-5


def test_scaled_and_integer_paths():
    y, u, v = noisy_clip_frame(48, 32)
    out = recon(y, u, v, width=80, height=56, bm=1)
    assert all(o.shape == (56, 80) and np.isfinite(o).all() for o in out)
    yi = np.full((32, 48), 64, np.uint8)
    ci = np.full((16, 24), 128, np.uint8)
    r = recon(yi, ci, ci, ss_w=1, ss_h=1, bm=1)
    assert [int(r[p][16, 24]) for p in range(3)] == [64, 128, 128]
    t = oracle.trecon_frame(oracle.trecon_params(bm=1), (yi, ci, ci), [(yi, ci, ci)], ss_w=1, ss_h=1)
    assert all(abs(int(t[p][16, 24]) - e) <= 1 for p, e in zip(range(3), (64, 128, 128)))
