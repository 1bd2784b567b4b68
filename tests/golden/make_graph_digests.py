#!/usr/bin/env python3
"""Writes tests/golden/graph_digests.json: the oracle's graphs of the
benchmark workloads at their full sizes, as digests (tests/digests.py).

  asm50k    50,000 contigs, S = 10 (the single-assembly GPU configuration)
  asm300k   300,000 contigs, S = 30 (the headline configuration)
  batch64   each of the 64 assemblies of the batched workload
  served    every request of the served workload: 24 sessions, each its
            initial assembly then 5 extensions of 2,000 contigs; request t of
            a session is the graph of its first t+1 pieces

Run from the repository root after __graft_entry__.build() (the oracle and
the synthetic generator are native libraries); it takes several minutes on
8 cores.  Usage: python tests/golden/make_graph_digests.py [names...]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(HERE))

from digests import input_digest, record_of_oracle  # noqa: E402
from oracle import oracle as orc  # noqa: E402
from rabbitbin_amd import synth  # noqa: E402

OUT = os.path.join(HERE, "graph_digests.json")


def concat(parts):
    seqs = np.concatenate([p[0] for p in parts])
    offs, base = [np.zeros(1, np.uint64)], 0
    for p in parts:
        offs.append(p[1][1:] - p[1][0] + np.uint64(base))
        base += int(p[1][-1] - p[1][0])
    depth = None if parts[0][2] is None else np.concatenate([p[2] for p in parts])
    return seqs, np.concatenate(offs), depth


def one(seqs, off, depth):
    S = 0 if depth is None else depth.shape[1]
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=1)
    r = record_of_oracle(ref, S)
    r["input"] = input_digest(seqs, off, depth)
    return r


def main(names):
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for name in names:
        t0 = time.time()
        if name in ("asm50k", "asm300k"):
            seqs, off, depth, _ = synth.workload(name)
            out[name] = one(seqs, off, depth)
        elif name == "batch64":
            asms, _ = synth.batch_workload(64)
            out[name] = [one(s, o, d) for s, o, d in asms]
        elif name == "served":
            _, pieces = synth.served_workload(24)
            out[name] = [[one(*concat(parts[:t + 1])) for t in range(len(parts))] for parts in pieces]
        else:
            raise SystemExit(f"unknown workload {name}")
        print(f"{name}: {time.time() - t0:.0f} s", flush=True)
        json.dump(out, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["asm50k", "asm300k", "batch64", "served"])
