"""The reference's gate study through the HIP backend: all 468 observations
(156 scenes x 3 degradations, 128x128 YUV420PS, jinc 3) in the study's four
Recon configurations, against the oracle sample by sample, and the published
figures of test/results/eval_gates_latest.md reproduced from the HIP
outputs."""
import numpy as np
import pytest

import eval_gates_study as study
import oracle
from lgcr_amd import abi

pytestmark = pytest.mark.gpu


def test_gate_study_through_hip(device):
    fmt = abi.make_format(np.float32, 128, 128)
    clips = {tag: abi.Clip(device, abi.default_params(kernel="jinc", taps=3, **kw), fmt, max_inflight=8)
             for tag, kw in study.CONFIGS}
    try:
        hip, obs = study.run(lambda tag, kw, planes: clips[tag].recon_host(*planes), threads=8)
    finally:
        for c in clips.values():
            c.close()
    ref, _ = study.run(lambda tag, kw, planes: oracle.recon_frame(
        oracle.make_params(128, 128, kernel="jinc", taps=3, **kw), *planes), obs, threads=8)
    mismatched, worst = 0, 0.0
    for key, outs in hip.items():
        for tag, planes in outs.items():
            for i in range(3):
                a, b = ref[key][tag][i], planes[i]
                d = float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64))))
                worst = max(worst, d)
                mismatched += int(np.count_nonzero(a != b))
    print(f"HIP vs oracle over {len(hip) * len(study.CONFIGS)} frames: {mismatched} samples differ, "
          f"max |d| {worst:g}")
    assert worst <= 1e-5
    got = study.summary(hip, obs)
    for key, want in study.PUBLISHED.items():
        assert got[key] == want, (key, got[key])
