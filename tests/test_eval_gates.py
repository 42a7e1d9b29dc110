"""Pins of the oracle's heuristic core against the reference's gate study
(test/eval_gates.py, results in test/results/eval_gates_latest.md).

* The study's scenes, regenerated here, hash-equal the study's own (hashes
  written by tests/golden/make_gate_fixtures.py with the study's functions).
* The oracle standing in for core.lgcr.Recon reproduces every figure of the
  published report that depends on the plugin's output: label counts and
  mean deltas of ungated algo 2 and algo 6, full algo 6 by degradation, the
  outcome counts and the eight worst residuals.  These run the jinc plain
  base, the guided kernel (algo 2, no gates), and algo 6 with and without
  its q and ms gates; a 1-2 % change of sigma, stretch or reg changes them.
* The mutual-structure gate agrees with the study's NumPy model ms_map
  (eval_gates.py:311-339), and the algo-6 affine credibility with its
  q_prod (eval_gates.py:356-393)."""
import os

import numpy as np
import pytest

import eval_gates_scene as scene
import eval_gates_study as study
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eval_gates.npz")


def oracle_recon(tag, kw, planes):
    return oracle.recon_frame(oracle.make_params(128, 128, kernel="jinc", taps=3, **kw), *planes)


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.fixture(scope="module")
def obs():
    return study.inputs()


def test_scenes_hash_equal_the_study(golden):
    import hashlib

    def sha(*arrays):
        h = hashlib.sha256()
        for a in arrays:
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    scenes = scene.scenes()
    assert [s[0] for s in scenes] == list(golden["names"])
    for si, (name, y, u, v, band) in enumerate(scenes):
        assert sha(y.astype(np.float32), u.astype(np.float32), v.astype(np.float32),
                   band.astype(np.uint8)) == golden["scene_sha256"][si], name
    for si, name, kind, y, u, v, band, u420, v420 in scene.observations():
        assert sha(u420.astype(np.float32), v420.astype(np.float32)) == \
            golden["chroma_sha256"][si][scene.KINDS.index(kind)], (name, kind)


def test_oracle_reproduces_published_study(obs):
    res, _ = study.run(oracle_recon, obs)
    got = study.summary(res, obs)
    for key, want in study.PUBLISHED.items():
        assert got[key] == want, (key, got[key])


@pytest.mark.parametrize("change", [{"sigma": 0.0102}, {"stretch": 1.02}, {"reg": 0.0051}, {"gsigma": 2.55}])
def test_study_figures_detect_small_parameter_changes(obs, change):
    """The published figures are precise enough to tell a 1-2 % change of a
    constant of the guided or affine arithmetic."""
    def recon(tag, kw, planes):
        return oracle_recon(tag, {**kw, **change}, planes)

    res, _ = study.run(recon, obs)
    got = study.summary(res, obs)
    assert any(got[k] != want for k, want in study.PUBLISHED.items())


def _observation(obs, key):
    si, kind = key.split("_")
    si = int(si[1:])
    return next(o for o in obs if o[0] == si and o[2] == kind)


def test_mutual_gate_matches_study_model(golden, obs):
    """or_mutual_gate (buildMutualGate, maps.cpp:831-948) on the footprint
    luma map vs the study's ms_map on the chroma samples of the band.  The
    model works in float64 without the plugin's 1e-12 / 1e-6 guards and
    clamps bilinear coordinates to size - 1.001, so values agree to float
    noise amplified by the smoothstep slopes (measured: max 1.6e-3, 0.2 % of
    samples above 1e-4); samples whose raw 3x3 tensor sum is below 1e-10,
    where the plugin returns 0 (maps.cpp:873-877) and the model does not,
    are left out.  A wrong gate constant moves gates by 0.1 or more."""
    L = oracle.lib()
    worst, over, n, total = 0.0, 0, 0, 0.0
    for key in golden["vector_keys"]:
        o = _observation(obs, key)
        y, u, v = o[3]
        lc = np.zeros((64, 64), np.float32)
        L.or_lc_map(oracle.fptr(y), 128, 128, 64, 64, 2.0, 2.0, -0.5, 0.0, oracle.fptr(lc))
        gate = np.zeros((64, 64), np.float32)
        L.or_mutual_gate(oracle.fptr(lc), oracle.fptr(u), oracle.fptr(v), 64, 64, None, 1, oracle.fptr(gate))
        bcm = golden[key + "_bcm"]
        keep = 9.0 * golden[key + "_trace"] >= 2e-10   # the model's trace is the tensor sum / 9
        d = np.abs(gate[bcm].astype(np.float64) - golden[key + "_ms"])[keep]
        if not d.size:   # flat luma over the whole band (the isoluminant scene)
            continue
        worst = max(worst, float(d.max()))
        over += int(np.count_nonzero(d > 1e-4))
        n += d.size
        total += float(d.sum())
    print(f"gate vs ms_map over {n} samples: max {worst:.3g}, mean {total / n:.3g}, above 1e-4: {over}")
    assert n > 30000
    assert worst <= 3e-3 and over <= 0.005 * n and total / n <= 2e-5


def test_affine_credibility_matches_study_model(golden, obs):
    """The algo-6 credibility g = mean(q) * (qmin / qmax) * varC / (varC +
    (sigma/4)^2) of buildAffineMaps (algos.cpp:1382-1526), dense path, vs the
    study's q_prod, at chroma samples at least 4 from the frame border.
    Nearer the border the plugin's polyphase tables shift the window start
    to stay inside the plane (kernels.cpp:149-150) where the model clamps
    sample indices, so the bilinear and bicubic candidates D(Y) differ
    there by construction."""
    L = oracle.lib()
    worst, n = 0.0, 0
    for key in golden["vector_keys"]:
        o = _observation(obs, key)
        g = np.zeros((64, 64), np.float32)
        L.or_affine_debug(oracle.fptr(g))
        try:
            oracle.recon_frame(oracle.make_params(128, 128, kernel="jinc", taps=3, algo=6, sparse=0), *o[3])
        finally:
            L.or_affine_debug(None)
        bcm = golden[key + "_bcm"]
        at = np.argwhere(bcm)
        inner = (at.min(axis=1) >= 4) & (at.max(axis=1) < 60)
        d = np.abs(g[bcm].astype(np.float64) - golden[key + "_qprod"])[inner]
        worst = max(worst, float(d.max()))
        n += d.size
    print(f"affine g vs q_prod over {n} interior samples: max {worst:.3g}")
    assert n > 30000 and worst <= 2e-6
