"""GPU: the benchmark workloads at their full sizes against the oracle,
through the digests tests/golden/make_graph_digests.py wrote from the oracle's
graphs (tests/golden/graph_digests.json): the 50,000- and 300,000-contig
assemblies, every assembly of the batched workload built as one batch
(captured and replayed), and every request of the served workload (Poisson
arrivals, five 2,000-contig extensions per session) served with continuous
batching.  A digest covers the sorted edge list (endpoints, sComp bits,
edgeScore bits), and the record also holds simCutoff, b0, the regime and
the evaluated pair count."""
import json
import os
import threading
import time

import numpy as np
import pytest

from digests import input_digest, record_of_result
from rabbitbin_amd import api, synth

pytestmark = pytest.mark.gpu

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "graph_digests.json")
DIG = json.load(open(PATH)) if os.path.exists(PATH) else {}


def _need(name):
    if name not in DIG:
        pytest.skip(f"no committed digest for {name}")
    return DIG[name]


def _same(want, g, what):
    got = record_of_result(g)
    for k in ("n_edges", "sim_cutoff", "b0", "regime", "gated_pairs", "digest"):
        assert got[k] == want[k], (what, k, got[k], want[k])


@pytest.mark.parametrize("name", ["asm50k", "asm300k"])
def test_single_assembly_digest(ctx, name):
    want = _need(name)
    seqs, off, depth, _ = synth.workload(name)
    assert input_digest(seqs, off, depth) == want["input"], "synthetic input differs on this host"
    asm = api.Assembly(ctx, seqs, off, depth)
    try:
        g = asm.build_graph(api.default_params(rand_seed=1))
    finally:
        asm.free()
    _same(want, g, name)


def test_batch64_digests(ctx):
    want = _need("batch64")
    asms, _ = synth.batch_workload(64)
    for k, (s, o, d) in enumerate(asms):
        assert input_digest(s, o, d) == want[k]["input"], ("synthetic input differs on this host", k)
    b = api.Batch(ctx, asms)
    try:
        p = api.default_params(rand_seed=1)
        b.build(p, capture=True)
        st = b.build(p, capture=True)
        assert st["replayed"]
        for k in range(len(asms)):
            _same(want[k], b.result(k), ("batch64", k))
    finally:
        b.free()


def _concat(parts):
    seqs = np.concatenate([p[0] for p in parts])
    offs, base = [np.zeros(1, np.uint64)], 0
    for p in parts:
        offs.append(p[1][1:] - p[1][0] + np.uint64(base))
        base += int(p[1][-1] - p[1][0])
    depth = None if parts[0][2] is None else np.concatenate([p[2] for p in parts])
    return seqs, np.concatenate(offs), depth


def test_served_workload_digests():
    """The Poisson arrival times are compressed 20x (80 requests/s), so that
    sessions overlap and requests share build steps."""
    want = _need("served")
    arrivals, pieces = synth.served_workload(24)
    arrivals = arrivals / 20.0
    for k, parts in enumerate(pieces):
        assert input_digest(*_concat(parts)) == want[k][-1]["input"], ("synthetic input differs", k)
    results = [[None] * len(p) for p in pieces]
    errors = []
    with api.Server(0) as srv:
        base = time.perf_counter()

        def client(k):
            try:
                d = base + arrivals[k] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ses, r = srv.submit(*pieces[k][0])
                results[k][0] = srv.wait(r, timeout=600)
                for t in range(1, len(pieces[k])):
                    results[k][t] = srv.wait(srv.extend(ses, *pieces[k][t]), timeout=600)
                srv.release(ses)
            except Exception as e:  # reported below
                errors.append((k, repr(e)))

        th = [threading.Thread(target=client, args=(k,)) for k in range(len(pieces))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    assert not errors, errors
    incr = 0
    for k, parts in enumerate(pieces):
        for t in range(len(parts)):
            g, info = results[k][t]
            assert info["status"] == 0
            incr += info["incremental"]
            _same(want[k][t], g, ("served", k, t))
    assert incr > 0  # extensions were served incrementally
    assert st["steps"] < st["requests"]  # requests shared steps
