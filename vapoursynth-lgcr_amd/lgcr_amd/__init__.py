"""MI355X-native LGCR chroma reconstruction: Python view of the HIP backend.

The compute lives in liblgcr_hip.so (csrc/); this package binds its C ABI
(include/lgcr_hip.h) and carries host utilities (build, synthetic clips).
"""
from . import paths  # noqa: F401
