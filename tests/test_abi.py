"""C ABI: the built library loads and exports every symbol the header
declares; host-side validation runs without a device."""
import ctypes as C
import os
import re

import pytest

from lgcr_amd import abi, paths

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "lgcr_hip.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(lgcr_\w+)\s*\(", text, re.M)))


def test_header_declares_binding_symbols():
    assert set(header_functions()) == set(abi.SYMBOLS)


def test_library_exports_every_symbol():
    lib = C.CDLL(paths.HIP_LIB)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_abi_version_and_defaults():
    assert abi.lib().lgcr_abi_version() == 1
    p = abi.default_params()
    # lgcrCreate defaults (plugin.cpp:1025-1133)
    assert (p.kernel, p.kp1, p.support, p.algo, p.strength, p.sparse) == (2, 3.0, 3.0, 2, 0.8, 1)
    assert (p.sigma, p.sratio, p.sdb, p.gsigma, p.rescue, p.ms, p.bp) == (0.01, 0.15, 3.0, 2.5, 1.0, 1.0, 0.0)
    assert (p.shift_x, p.shift_y, p.ar_margin, p.ridge, p.cedge, p.bm) == (-0.5, 0.0, 0.0, 1, 0, 0)


def test_no_device_is_a_clean_error(monkeypatch):
    # on a host without a HIP device the backend refuses with a status and a
    # message, it never falls back to CPU compute
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    with pytest.raises(abi.LgcrError) as e:
        abi.Device(0)
    assert e.value.status == abi.LGCR_INVALID_CONTEXT
    assert "no HIP device" in str(e.value)
