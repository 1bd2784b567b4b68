#!/usr/bin/env python3
"""Write tests/golden/probmh_golden.npz from oracle/_ref/libref_probmh.so.

oracle/_ref is the reference's own src/probmh.cpp (with src/probmh.h and
src/hash_int.h) compiled unchanged by oracle/Makefile (`make ref`); the
fixtures below are its outputs, so the restatement in oracle/ and the
product's host tables can be checked against them where the reference is not
present.  Run from the repository root:  python tests/golden/make_golden.py

Contents (all arrays; the file holds no pickled objects):
  fmix_in [64,2] u64, fmix_out [64] u64        mc::murmur3_fmix  (hash_int.h:79-87)
  wy_states [S] u64, wyrand [S,32] u64         wyrand_next       (probmh.cpp:60-87)
  uni_bounds [B] u32, uni [S,B,32] u32         wy_uniform_int
  zig [2,512] f64                              zig_exponential   (probmh.cpp:319-338)
  zig_x, zig_y [257] f64                       ziggurat tables   (probmh.cpp:96-314)
  ted_c [5,3] f64, ted [5,64] f64              ted_sample        (probmh.cpp:344-357)
  tw_m{2,16}_T [136,m] f64, _W [136,m] u64     a fresh sketch after addHash(x, 1.0),
                                               x = each canonical 4-mer code
  tw_m500_sha256_T, _W  (u8 digests)           the same at m = 500, as SHA-256 of the bytes
  sk_m{2,16,500}_codes/_w [V,136]              weight vectors (0 = element absent)
  sk_m{..}_win [V,m] u64, _reg [V,m] f64       streaming sketch after inserting the
                                               present elements in ascending code order
"""
import hashlib
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, REPO)
from oracle import oracle as orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probmh_golden.npz")


def canonical_codes():
    out = []
    for v in range(256):
        rc = 0
        for d in range(4):
            rc = (rc << 2) | (3 - ((v >> (2 * d)) & 3))
        if v <= rc:
            out.append(v)
    return np.array(out, np.uint64)


def main():
    if not orc.ref_available():
        sys.exit("oracle/_ref/libref_probmh.so is missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20261203)
    g = {}
    fin = np.stack([rng.integers(0, 2**63, 64, dtype=np.uint64) * np.uint64(2) + np.uint64(1),
                    np.array([42] * 32 + list(rng.integers(0, 2**63, 32, dtype=np.uint64)), np.uint64)], 1)
    fin[0] = [0, 42]
    fin[1] = [255, 42]
    g["fmix_in"] = fin
    g["fmix_out"] = np.array([orc.fmix64(int(a), int(b), "ref") for a, b in fin], np.uint64)

    states = np.array([0, 1, 0xDEADBEEF, 2**64 - 1, 0x9E3779B97F4A7C15], np.uint64)
    g["wy_states"] = states
    g["wyrand"] = np.stack([orc.wyrand_seq(int(s), 32, "ref")[0] for s in states])
    bounds = np.array([1, 2, 3, 136, 499, 500, 2**31 + 7, 2**32 - 1], np.uint32)
    g["uni_bounds"] = bounds
    g["uni"] = np.stack([np.stack([orc.wy_uniform_int_seq(int(s), int(b), 32, "ref")[0] for b in bounds])
                         for s in states])
    g["zig"] = np.stack([orc.zig_seq(int(s), 512, "ref")[0] for s in (7, 0xC0FFEE)])
    g["zig_x"], g["zig_y"] = orc.zig_tables_ref()
    ted, _ = orc.ted_params(500)
    rows = [0, 1, 10, 100, 497]
    g["ted_c"] = ted[rows][:, 2:5].copy()
    g["ted"] = np.stack([orc.ted_seq(*ted[r, 2:5], 1234 + r, 64, "ref")[0] for r in rows])

    codes = canonical_codes()
    for m in (2, 16, 500):
        T = np.zeros((136, m))
        W = np.zeros((136, m), np.uint64)
        for x, v in enumerate(codes):
            W[x], T[x] = orc.pmh_stream([v], [1.0], m, which="ref")
        if m == 500:
            g["tw_m500_sha256_T"] = np.frombuffer(hashlib.sha256(T.tobytes()).digest(), np.uint8)
            g["tw_m500_sha256_W"] = np.frombuffer(hashlib.sha256(W.tobytes()).digest(), np.uint8)
        else:
            g[f"tw_m{m}_T"], g[f"tw_m{m}_W"] = T, W
        V = {2: 8, 16: 20, 500: 4}[m]
        wts = np.zeros((V, 136))
        for t in range(V):
            present = rng.random(136) < [0.9, 0.3, 0.05, 1.0][t % 4]
            w = np.exp(rng.normal(0, 1.5, 136))
            if t % 5 == 3:
                w = np.full(136, 1.0)  # equal weights: ties decided by proposal order
            if t % 7 == 6:
                w = np.minimum(w * 30, 20.0)  # the reference's weight cap
            wts[t] = np.where(present, w, 0.0)
        wins = np.zeros((V, m), np.uint64)
        regs = np.zeros((V, m))
        for t in range(V):
            sel = wts[t] > 0
            wins[t], regs[t] = orc.pmh_stream(codes[sel], wts[t][sel], m, which="ref")
        g[f"sk_m{m}_w"] = wts
        g[f"sk_m{m}_win"] = wins
        g[f"sk_m{m}_reg"] = regs
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
