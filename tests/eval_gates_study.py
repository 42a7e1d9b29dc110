"""The Recon side of the reference's gate study (test/eval_gates.py:441-500):
every observation reconstructed with jinc 3 in four configurations, the mean
absolute chroma error over the evaluation band, and the summary figures the
study prints.  `recon` is the implementation under test (the oracle on the
CPU, the HIP backend on the GPU)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import eval_gates_scene as scene

# (tag, Recon arguments) as the study runs them (eval_gates.py:460-466)
CONFIGS = (
    ("plain", dict(strength=0.0)),
    ("algo2_ungated", dict(strength=0.8, algo=2, ms=0.0, cedge=0)),
    ("algo6_ungated", dict(strength=0.8, algo=6, qgate=0.0, ms=0.0)),
    ("algo6_full", dict(strength=0.8, algo=6, qgate=1.0, ms=1.0)),
)

# test/results/eval_gates_latest.md: the figures that depend on the plugin's
# output (the statistics columns come from the study's own NumPy maps)
PUBLISHED = {
    "delta2": "468 observations, 420 labeled (135 benefit / 285 harm), 48 neutral excluded; mean delta=+0.00263",
    "delta6": "468 observations, 441 labeled (205 benefit / 236 harm), 27 neutral excluded; mean delta=+0.00003",
    "gain6_by_d": ["  box  -0.00189", "  tri  -0.00150", "  bic  -0.00026", "  all  -0.00122"],
    "outcomes": "full algo=6 labeled outcomes: 158 benefit / 88 harm / 222 neutral",
    "worst": [
        ("edge_s5_m0.5_a45_c0.55", "bic", "+0.00844"),
        ("edge_s5_m0_a45_c0.55", "bic", "+0.00765"),
        ("edge_s5_m0_a22_c0.55", "bic", "+0.00737"),
        ("edge_s5_m0.5_a22_c0.55", "bic", "+0.00648"),
        ("edge_s3_m1_a0_c0.55", "bic", "+0.00514"),
        ("edge_s5_m0.5_a0_c0.55", "bic", "+0.00459"),
        ("edge_s5_m1_a45_c0.55", "bic", "+0.00454"),
        ("edge_s5_m0_a45_c0.55", "tri", "+0.00444"),
    ],
}


def inputs():
    """[(scene index, name, kind, (Y, U420, V420) float32, U, V float64, band)]"""
    out = []
    for si, name, kind, y, u, v, band, u420, v420 in scene.observations():
        out.append((si, name, kind, (y, u420.astype(np.float32), v420.astype(np.float32)),
                    u.astype(np.float64), v.astype(np.float64), band))
    return out


def band_mae(planes, u64, v64, band):
    ou, ov = planes[1].astype(np.float64), planes[2].astype(np.float64)
    return float((np.abs(ou - u64)[band].mean() + np.abs(ov - v64)[band].mean()) / 2)


def run(recon, obs=None, threads=8):
    """recon(tag, kwargs, (Y, U420, V420)) -> output planes.  Returns
    {(scene index, kind): {tag: output planes}} and the inputs."""
    obs = inputs() if obs is None else obs
    jobs = [(o, tag, kw) for o in obs for tag, kw in CONFIGS]

    def one(job):
        o, tag, kw = job
        return (o[0], o[2]), tag, recon(tag, kw, o[3])

    res = {}
    with ThreadPoolExecutor(threads) as ex:
        for key, tag, planes in ex.map(one, jobs):
            res.setdefault(key, {})[tag] = planes
    return res, obs


def summary(res, obs):
    """The plugin-dependent lines of the study's report (eval_gates.py:486-500,
    report_classifier 406-412)."""
    rows = []
    for si, name, kind, planes, u64, v64, band in obs:
        mae = {tag: band_mae(res[(si, kind)][tag], u64, v64, band) for tag, _ in CONFIGS}
        rows.append({"name": name, "d": kind, "delta2": mae["algo2_ungated"] - mae["plain"],
                     "delta6": mae["algo6_ungated"] - mae["plain"], "gain6": mae["algo6_full"] - mae["plain"]})

    def label_line(key, eps=5e-4):
        lab = [r for r in rows if abs(r[key]) > eps]
        return (f"{len(rows)} observations, {len(lab)} labeled ({sum(r[key] < 0 for r in lab)} benefit / "
                f"{sum(r[key] > 0 for r in lab)} harm), {len(rows) - len(lab)} neutral excluded; "
                f"mean delta={np.mean([r[key] for r in rows]):+.5f}")

    full = [r for r in rows if abs(r["gain6"]) > 5e-4]
    return {
        "delta2": label_line("delta2"),
        "delta6": label_line("delta6"),
        "gain6_by_d": [f"  {k:<4} {np.mean([r['gain6'] for r in rows if r['d'] == k]):+.5f}" for k in scene.KINDS]
        + [f"  all  {np.mean([r['gain6'] for r in rows]):+.5f}"],
        "outcomes": (f"full algo=6 labeled outcomes: {sum(r['gain6'] < 0 for r in full)} benefit / "
                     f"{sum(r['gain6'] > 0 for r in full)} harm / {len(rows) - len(full)} neutral"),
        "worst": [(r["name"], r["d"], f"{r['gain6']:+.5f}")
                  for r in sorted(rows, key=lambda r: r["gain6"], reverse=True)[:8]],
    }
