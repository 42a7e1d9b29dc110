"""In-tree build of the native libraries (hipcc for gfx950, gcc for host-only C)."""
from __future__ import annotations

import os
import subprocess

from . import paths

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: every product/sum rounds as written, like the scalar
# reference; packed FP32 ops are disabled for correctness on this platform.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]
HIP_SOURCES = ["lgcr_kernels.hip", "lgcr_runtime.cpp"]


def build_hip(verbose=False):
    inc = os.path.join(os.path.dirname(paths.ROOT_DIR), "include")
    srcs = [os.path.join(paths.CSRC_DIR, s) for s in HIP_SOURCES]
    cmd = [HIPCC, f"--offload-arch={ARCH}", *HIP_FLAGS, f"-I{inc}", f"-I{paths.CSRC_DIR}", *srcs,
           "-o", paths.HIP_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(paths.HIP_LIB + ".tmp", paths.HIP_LIB)
    if verbose:
        print(" ".join(cmd))
    return paths.HIP_LIB


def build_synth():
    cmd = ["gcc", "-O3", "-Wall", "-fPIC", "-shared", "-pthread", os.path.join(paths.CSRC_DIR, "synth.c"),
           "-o", paths.SYNTH_LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(paths.SYNTH_LIB + ".tmp", paths.SYNTH_LIB)
    return paths.SYNTH_LIB


def build_vsplugin():
    """The VapourSynth 4 plugin glue (liblgcr_vs.so, next to liblgcr_hip.so,
    which it links with an $ORIGIN rpath).  It holds no device code."""
    inc = os.path.join(os.path.dirname(paths.ROOT_DIR), "include")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", f"-I{os.path.join(paths.CSRC_DIR, 'vs4')}",
           f"-I{inc}", os.path.join(paths.CSRC_DIR, "lgcr_vsplugin.cpp"), f"-L{paths.PKG_DIR}", "-llgcr_hip",
           "-Wl,-rpath,$ORIGIN", "-o", paths.VS_PLUGIN + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(paths.VS_PLUGIN + ".tmp", paths.VS_PLUGIN)
    return paths.VS_PLUGIN


def build_all(verbose=False):
    build_synth()
    build_hip(verbose)
    build_vsplugin()
