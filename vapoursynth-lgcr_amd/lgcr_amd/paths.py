"""Locations of the in-tree built libraries."""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
CSRC_DIR = os.path.join(ROOT_DIR, "csrc")
HIP_LIB = os.path.join(PKG_DIR, "liblgcr_hip.so")
SYNTH_LIB = os.path.join(PKG_DIR, "liblgcr_synth.so")
VS_PLUGIN = os.path.join(PKG_DIR, "liblgcr_vs.so")   # VapourSynth 4 plugin (VapourSynthPluginInit2)
