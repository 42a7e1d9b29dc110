"""The reference's quality battery through the HIP backend: every Recon and
TRecon frame of the battery (the 14 cases in four configurations, the
temporal pair, the TRecon cases and the integer round trip) against the
oracle sample by sample, and the published figures of
test/results/latest.md reproduced from the HIP outputs (the same entries
and bounds as the oracle's own check in test_battery.py)."""
import numpy as np
import pytest

import battery_published as pub
from lgcr_amd import battery

pytestmark = pytest.mark.gpu


def test_battery_through_hip(device):
    hip_recon, hip_trecon, close = battery.hip_runners(device)
    stats = {"frames": 0, "differ": 0, "worst": 0.0}

    def compare(got, ref):
        stats["frames"] += 1
        for a, b in zip(ref, got):
            d = float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))
            stats["worst"] = max(stats["worst"], d)
            stats["differ"] += int(np.count_nonzero(a != b))

    def recon(kw, planes, ow, oh):
        got = hip_recon(kw, planes, ow, oh)
        compare(got, pub.oracle_recon(kw, planes, ow, oh))
        return got

    def trecon(kw, center, nbrs):
        got = hip_trecon(kw, center, nbrs)
        compare(got, pub.oracle_trecon(kw, center, nbrs))
        return got

    try:
        case_res, temporal, trc, int8 = battery.run_all(recon, trecon, all_algos=True)
    finally:
        close()
    print(f"HIP vs oracle over {stats['frames']} battery frames: {stats['differ']} samples differ, "
          f"max |d| {stats['worst']:g}")
    assert stats["worst"] <= 1e-5
    for name, want in pub.TABLE.items():
        if name == "upscale2x":
            continue
        for tag, w in zip(pub.TAGS, want):
            bound = pub.DEVIATIONS.get((name, tag))
            got = case_res[name][tag]
            if bound is None:
                assert f"{got:.5f}" == f"{w:.5f}", (name, tag, got)
            else:
                assert abs(got - w) <= bound, (name, tag, got)
    for label, want in pub.TRECON.items():
        for tag, w in zip(("plain", "algo2", "trecon"), want):
            assert f"{trc[label][tag]:.5f}" == f"{w:.5f}", (label, tag)
    for tag, w in zip(pub.TAGS, pub.TEMPORAL):
        assert f"{temporal[tag]:.5f}" == f"{w:.5f}", tag
    assert int8 == (pub.INT8, True)
    battery.check_gates(case_res, trc, int8[1])
