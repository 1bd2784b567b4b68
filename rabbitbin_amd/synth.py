"""Seeded synthetic assemblies (librbg_synth.so) for the benchmark workloads.

Workloads (master seed 20261203, sub-seed = workload index):
  asm2k     2,000 contigs, 20 genomes, lengths LogNormal(ln 8000, 0.8), 3 samples
  asm50k   50,000 contigs, 200 genomes, LogNormal(ln 8000, 0.8), 10 samples
  asm300k 300,000 contigs, 1,000 genomes, LogNormal(ln 6000, 1.0), 30 samples
  batch64  64 assemblies, contigs log-uniform in [500, 40,000], 50-400 genomes
           each, samples from {3, 10, 30}, lengths LogNormal(ln 8000, 0.8)
  served   Poisson arrivals (mean 4/s) of assemblies with 1,000-100,000
           contigs (log-uniform), 3-30 samples, each followed by 5 batches of
           2,000 more contigs of the same genomes
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


def batch_workload(count=64, threads=0):
    """The batched workload: list of (seqs, offsets, depth) and the specs
    (n, genomes, S) of its assemblies."""
    rng = np.random.default_rng(MASTER_SEED * 16 + 3)
    ns = np.exp(rng.uniform(np.log(500), np.log(40000), count)).astype(np.int64)
    gs = rng.integers(50, 401, count)
    Ss = rng.choice(np.array([3, 10, 30]), count)
    out, spec = [], []
    for k in range(count):
        seqs, off, depth, _ = make_assembly(int(ns[k]), int(gs[k]), math.log(8000), 0.8, int(Ss[k]),
                                            seed=(MASTER_SEED * 16 + 3) * 1000 + k, threads=threads)
        out.append((seqs, off, depth))
        spec.append((int(ns[k]), int(gs[k]), int(Ss[k])))
    return out, spec


def split_assembly(seqs, offsets, depth, sizes):
    """Consecutive pieces of an assembly (contig counts `sizes`), each as
    (seqs, offsets from 0, depth)."""
    parts, c = [], 0
    for m in sizes:
        o = offsets[c:c + m + 1]
        b0, b1 = int(o[0]), int(o[-1])
        parts.append((seqs[b0:b1], (o - o[0]).astype(np.uint64), None if depth is None else depth[c:c + m]))
        c += m
    return parts


def served_workload(count=24, batches=5, batch_size=2000, rate=4.0, threads=0):
    """The served workload: arrival times (s) and, per assembly, its pieces:
    the initial assembly then `batches` extensions of `batch_size` contigs."""
    rng = np.random.default_rng(MASTER_SEED * 16 + 4)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    arrivals -= arrivals[0]
    ns = np.exp(rng.uniform(np.log(1000), np.log(100000), count)).astype(np.int64)
    Ss = rng.integers(3, 31, count)
    out = []
    for k in range(count):
        n = int(ns[k])
        g = int(min(1000, max(10, n // 100)))
        total = n + batches * batch_size
        seqs, off, depth, _ = make_assembly(total, g, math.log(8000), 0.8, int(Ss[k]),
                                            seed=(MASTER_SEED * 16 + 4) * 1000 + k, threads=threads)
        out.append(split_assembly(seqs, off, depth, [n] + [batch_size] * batches))
    return arrivals, out
