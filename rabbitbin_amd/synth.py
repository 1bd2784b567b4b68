"""Seeded synthetic assemblies (librbg_synth.so) for the benchmark workloads.

Workloads (master seed 20261203, sub-seed = workload index):
  smoke     2,000 contigs, 20 genomes, lengths LogNormal(ln 8000, 0.8), 3 samples
  asm50k   50,000 contigs, 200 genomes, LogNormal(ln 8000, 0.8), 10 samples
  asm300k 300,000 contigs, 1,000 genomes, LogNormal(ln 6000, 1.0), 30 samples
"""
import ctypes as C
import math

import numpy as np

from . import SYNTH_LIB_PATH

MASTER_SEED = 20261203

WORKLOADS = {
    "asm2k": dict(n=2000, genomes=20, mu=math.log(8000), sigma=0.8, S=3, sub=0),
    "asm50k": dict(n=50000, genomes=200, mu=math.log(8000), sigma=0.8, S=10, sub=1),
    "asm300k": dict(n=300000, genomes=1000, mu=math.log(6000), sigma=1.0, S=30, sub=2),
}

_lib = None


def _synth():
    global _lib
    if _lib is None:
        L = C.CDLL(SYNTH_LIB_PATH)
        L.rbg_synth_layout.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                       C.c_void_p, C.c_void_p]
        L.rbg_synth_fill.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_uint32, C.c_void_p, C.c_int]
        _lib = L
    return _lib


def make_assembly(n, genomes, mu, sigma, S, seed, threads=0):
    """Returns (seqs uint8[total], offsets uint64[n+1], depth float32[n,S] or None, gid uint32[n])."""
    L = _synth()
    lengths = np.zeros(n, np.uint32)
    gid = np.zeros(n, np.uint32)
    L.rbg_synth_layout(seed, n, genomes, mu, sigma, lengths.ctypes.data, gid.ctypes.data)
    offsets = np.zeros(n + 1, np.uint64)
    np.cumsum(lengths, out=offsets[1:])
    seqs = np.empty(int(offsets[-1]), np.uint8)
    depth = np.zeros((n, S), np.float32) if S else None
    L.rbg_synth_fill(seed, n, genomes, lengths.ctypes.data, gid.ctypes.data, offsets.ctypes.data,
                     seqs.ctypes.data, S, depth.ctypes.data if S else None, threads)
    return seqs, offsets, depth, gid


def workload(name, threads=0, n=None):
    w = dict(WORKLOADS[name])
    if n is not None:
        w["n"] = n
    seed = MASTER_SEED * 16 + w["sub"]
    return make_assembly(w["n"], w["genomes"], w["mu"], w["sigma"], w["S"], seed, threads)
