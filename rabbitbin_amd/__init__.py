"""rabbitbin_amd: MI355X (gfx950) implementation of RabbitBin's graph-construction path.

The compute runs in ``librbg.so`` (hand-written HIP kernels behind the C ABI of
``include/rbg.h``); this package only binds that ABI with ctypes.  Importing
``rabbitbin_amd.api`` raises if the native library has not been built: there is
no CPU fallback.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "librbg.so")
SYNTH_LIB_PATH = os.path.join(PKG_DIR, "librbg_synth.so")

__all__ = ["PKG_DIR", "REPO_DIR", "LIB_PATH", "SYNTH_LIB_PATH"]
