"""lgcr_downsample_frame_noalloc (intermediates in a fixed per-GPU array,
no allocation) against lgcr_downsample_frame on the same device frames:
two calls in a row, the second smaller, both bit-identical; frames beyond
the array's size and host frames are refused."""
import numpy as np
import pytest

from lgcr_amd import abi, synth

pytestmark = pytest.mark.gpu


def to_dev(planes):
    import torch
    return [torch.from_numpy(p.view(np.int16) if p.dtype == np.uint16 else p).to("cuda") for p in planes]


def planes_of(ts):
    return [(t.data_ptr(), t.stride(0) * t.element_size()) for t in ts]


def run_both(device, w, h, bits, **kw):
    import torch
    planes = synth.clip_frame(w, h, ss_w=0, ss_h=0, bits=bits, seed=w + h, polygons=16)
    dt = planes[0].dtype
    fmt = abi.make_format(dt, w, h, 0, 0, bits if bits == 10 else None)
    din = to_dev(planes)
    tdt = torch.float32 if dt == np.float32 else (torch.uint8 if dt == np.uint8 else torch.int16)
    outs = {}
    clip = abi.DownsampleClip(device, abi.downsample_params(**kw), fmt)
    try:
        for mode in ("workspace", "noalloc"):
            dout = [torch.full((h // 2, w // 2), 7, dtype=tdt, device="cuda") for _ in range(2)]
            args = (planes_of(din), [(0, 0)] + planes_of(dout))
            if mode == "noalloc":
                clip.downsample_device_noalloc(*args)
            else:
                clip.downsample_device(*args)
            torch.cuda.synchronize()
            outs[mode] = [t.cpu().numpy() for t in dout]
    finally:
        clip.close()
    return outs


@pytest.mark.parametrize("quality", [2, 0])
def test_noalloc_matches_workspace_path_twice(device, quality):
    # first call at 1920x1080 P10, second smaller (640x360 P8): the second
    # reuses the array the first left behind
    for w, h, bits in ((1920, 1080, 10), (640, 360, 8)):
        outs = run_both(device, w, h, bits, quality=quality, kernel="spline36")
        for a, b in zip(outs["workspace"], outs["noalloc"]):
            assert a.shape == b.shape and np.array_equal(a, b), (w, h, quality)


def test_noalloc_float_quality1(device):
    outs = run_both(device, 256, 144, 32, quality=1, kernel="lanczos3", loc="topleft")
    for a, b in zip(outs["workspace"], outs["noalloc"]):
        assert np.array_equal(a, b)


def test_noalloc_refuses_oversize_and_host_frames(device):
    import torch
    w, h = 3840, 2160
    fmt = abi.make_format(np.uint16, w, h, 0, 0, 10)
    clip = abi.DownsampleClip(device, abi.downsample_params(quality=2), fmt, max_inflight=1)
    try:
        din = [torch.zeros((h, w), dtype=torch.int16, device="cuda") for _ in range(3)]
        dout = [torch.zeros((h // 2, w // 2), dtype=torch.int16, device="cuda") for _ in range(2)]
        with pytest.raises(abi.LgcrError) as e:
            clip.downsample_device_noalloc(planes_of(din), [(0, 0)] + planes_of(dout))
        assert e.value.status == abi.LGCR_UNSUPPORTED and "scratch" in str(e.value)
    finally:
        clip.close()
    small = abi.DownsampleClip(device, abi.downsample_params(quality=0), abi.make_format(np.uint8, 64, 32, 0, 0))
    try:
        y = np.zeros((32, 64), np.uint8)
        fin, keep = abi.host_frame((y, y, y))
        o = np.zeros((16, 32), np.uint8)
        fout, keep2 = abi.host_frame((y, o, o))
        import ctypes as C
        st = abi.lib().lgcr_downsample_frame_noalloc(small.handle, C.byref(fin), C.byref(fout))
        assert st == abi.LGCR_INVALID_ARGUMENT
    finally:
        small.close()
