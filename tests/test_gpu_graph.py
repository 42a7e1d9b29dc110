"""Steady-state frame graphs: device-resident Recon frames replay one
captured hipGraph per workspace, the frames' planes reaching it through a
device-side table.  Every replay must equal the direct launches and the
oracle; replays are interleaved with other device work and host
synchronisation, and the graph is re-captured when the siting changes."""
import numpy as np
import pytest
import torch

import oracle
from lgcr_amd import abi
from parity import compare_planes, frame

pytestmark = pytest.mark.gpu

CASES = [
    # (id, width, height, ss_w, ss_h, bits, params)
    ("p10-420-bp", 320, 180, 1, 1, 10, {"bp": 0.5}),
    ("p8-422", 160, 96, 1, 0, 8, {}),
    ("ps-444", 128, 64, 0, 0, 32, {}),
    ("p8-420-upscale", 160, 96, 1, 1, 8, {"width": 240, "height": 144}),
    ("p10-420-algo4", 160, 96, 1, 1, 10, {"algo": 4}),
]


def _dev(a):
    t = torch.from_numpy(a.view(np.int16) if a.dtype == np.uint16 else a)
    return t.to("cuda")


def _planes(frames_dev, bps):
    return [[(t.data_ptr(), t.stride(0) * bps) for t in f] for f in frames_dev]


@pytest.fixture
def no_capture(device, monkeypatch):
    monkeypatch.setenv("LGCR_CAPTURE", "0")
    return device


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_graph_replays_equal_direct_and_oracle(no_capture, case):
    tag, w, h, ssw, ssh, bits, params = case
    n = 7
    frames = [frame(w, h, ss_w=ssw, ss_h=ssh, bits=bits, seed=21, frame=k, polygons=10) for k in range(n)]
    dt = frames[0][0].dtype
    bps = dt.itemsize
    ow, oh = params.get("width", w), params.get("height", h)
    fmt = abi.make_format(dt, w, h, ssw, ssh, None if bits in (8, 16, 32) else bits)
    clip = abi.Clip(no_capture, abi.default_params(**params), fmt, max_inflight=3)
    din = [[_dev(p) for p in f] for f in frames]
    tdt = {1: torch.uint8, 2: torch.int16, 4: torch.float32}[bps]
    same = (ow, oh) == (w, h)

    def run(graphs):
        clip.set_graphs(graphs)
        outs = [[torch.full((oh, ow), 7, dtype=tdt, device="cuda") for _ in range(3)] for _ in range(n)]
        fin = _planes(din, bps)
        fout = [[(0, 0) if same else (o[0].data_ptr(), o[0].stride(0) * bps)] +
                [(t.data_ptr(), t.stride(0) * bps) for t in o[1:]] for o in outs]
        clip.recon_batch(fin, fout)
        torch.cuda.synchronize()
        return [[(t.cpu().numpy().view(dt) if bps == 2 else t.cpu().numpy()) for t in o] for o in outs]

    try:
        direct = run(False)
        for rep in range(3):
            got = run(True)
            # other device work and host synchronisation between the replays
            clip.set_graphs(False)
            clip.recon_host(*frames[rep])
            x = torch.randn(1 << 20, device="cuda")
            (x * 2).sum().item()
            torch.cuda.synchronize()
            for k in range(n):
                for i in (range(3) if not same else (1, 2)):
                    assert np.array_equal(got[k][i], direct[k][i]), f"{tag} replay {rep} frame {k} plane {i}"
        # the graph result against the oracle
        op = {k: v for k, v in params.items() if k not in ("width", "height")}
        for k in (0, n - 1):
            ref = oracle.recon_frame(oracle.make_params(ow, oh, **op), *frames[k],
                                     bits=None if bits in (8, 16, 32) else bits, ss_w=ssw, ss_h=ssh)
            for i in (range(3) if not same else (1, 2)):
                cnt, mx = compare_planes(ref[i], got[k][i], bits == 32)
                print(f"{tag} frame {k} plane {i}: mismatches={cnt} max_abs={mx:g}")
                assert cnt == 0
    finally:
        clip.close()


def test_graph_recaptured_for_new_siting(no_capture):
    """Frames of another siting use another geometry: the workspace graph is
    captured again and the results follow the new siting."""
    w, h = 192, 108
    frames = [frame(w, h, bits=10, seed=4, frame=k) for k in range(4)]
    fmt = abi.make_format(np.uint16, w, h, bits=10)
    din = [[_dev(p) for p in f] for f in frames]
    clip = abi.Clip(no_capture, abi.default_params(), fmt, max_inflight=2)
    try:
        for sx, sy in ((-0.5, 0.0), (0.0, -0.5), (-0.5, 0.0)):
            outs = [[torch.empty((h, w), dtype=torch.int16, device="cuda") for _ in range(2)] for _ in range(4)]
            for k in range(4):
                planes_in = [(t.data_ptr(), t.stride(0) * 2) for t in din[k]]
                planes_out = [(0, 0)] + [(t.data_ptr(), t.stride(0) * 2) for t in outs[k]]
                clip.recon_device(planes_in, planes_out, shift_x=sx, shift_y=sy)
            torch.cuda.synchronize()
            for k in (0, 3):
                ref = oracle.recon_frame(oracle.make_params(w, h, shift_x=sx, shift_y=sy), *frames[k], bits=10)
                for i in (1, 2):
                    got = outs[k][i - 1].cpu().numpy().view(np.uint16)
                    cnt, mx = compare_planes(ref[i], got, False)
                    assert cnt == 0, f"siting ({sx}, {sy}) frame {k} plane {i}: {cnt} mismatches"
    finally:
        clip.close()
