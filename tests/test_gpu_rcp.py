"""The reciprocal premise behind bit-exact tap weights: on the device, the
hardware reciprocal refined by one FMA Newton step (used for 1/y, y >= 1, in
the guided tap weights and the Downsample spatial weights) equals the IEEE
quotient 1.0f / y for every float in [1, 2^126)."""
import struct

import pytest

pytestmark = pytest.mark.gpu


def fbits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_rcp_nr_equals_division_exhaustive(device):
    from lgcr_amd import abi
    lo, hi = fbits(1.0), fbits(2.0 ** 126)
    assert hi - lo == 126 << 23   # 126 binades, 2^23 floats each
    n, first = abi.check_rcp(0, lo, hi)
    assert n == 0, f"{n} mismatches, first at bit pattern {first:#010x}"

