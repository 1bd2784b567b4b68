"""GPU: the batched form.  Every assembly of a batch must come out exactly as
its plain (single-assembly) build, whatever the work schedule, and replays of
a captured build must equal the eager build."""
import numpy as np
import pytest

from rabbitbin_amd import api, synth

pytestmark = pytest.mark.gpu


def _mixed_assemblies():
    # (n, genomes, S): both regimes (26,000 > 25,000 is fused), depth-less and
    # one-sample assemblies, and S on both sides of the 16- and 32-wide chunks
    spec = [(700, 8, 3, 1), (1500, 20, 10, 2), (2600, 30, 30, 3), (900, 10, 0, 4), (1200, 12, 1, 5),
            (26000, 200, 10, 6), (3000, 25, 40, 7)]
    out = []
    for n, g, S, seed in spec:
        seqs, off, depth, _ = synth.make_assembly(n, g, np.log(8000), 0.8, S, seed=1000 + seed)
        out.append((seqs, off, depth if S else None))
    return out


def _same(r, q):
    assert r["regime"] == q["regime"] and r["b0"] == q["b0"] and r["sim_cutoff"] == q["sim_cutoff"]
    assert r["gated_pairs"] == q["gated_pairs"]
    for k in ("from", "to", "scomp"):
        assert np.array_equal(r[k], q[k]), k
    assert (r["score"] is None) == (q["score"] is None)
    if r["score"] is not None:
        assert np.array_equal(r["score"], q["score"])


@pytest.fixture(scope="module")
def mixed():
    return _mixed_assemblies()


@pytest.fixture(scope="module")
def plain(ctx, mixed):
    p = api.default_params(rand_seed=1)
    return [ctx.build_graph(s, o, d, p) for s, o, d in mixed]


@pytest.mark.parametrize("tiles,rows", [(16, 64), (3, 17), (32, 128), (64, 4096)])
def test_batch_equals_plain(ctx, mixed, plain, tiles, rows):
    b = api.Batch(ctx, mixed)
    try:
        b.set_schedule(tiles, rows)
        st = b.build(api.default_params(rand_seed=1))
        assert st["edges"] == sum(len(p["from"]) for p in plain)
        for k in range(len(mixed)):
            _same(b.result(k), plain[k])
    finally:
        b.free()


def test_batch_graph_replays_equal_eager(ctx, mixed, plain):
    b = api.Batch(ctx, mixed)
    try:
        p = api.default_params(rand_seed=1)
        first = b.build(p, capture=True)
        assert not first["replayed"]
        for rep in range(3):
            b.build(p)  # eager build between replays
            st = b.build(p, capture=True)
            assert st["replayed"]
            for k in range(len(mixed)):
                _same(b.result(k), plain[k])
    finally:
        b.free()


def test_single_assembly_capture(ctx):
    seqs, off, depth, _ = synth.workload("asm50k", n=26000)
    p = api.default_params(rand_seed=1)
    asm = api.Assembly(ctx, seqs, off, depth)
    try:
        ref = asm.build_graph(p)
        for rep in range(3):
            r = asm.build_graph(p, capture=True)
            _same(r, ref)
            asm.build_graph(p)
    finally:
        asm.free()


def test_autotuned_batch_equals_plain(ctx, mixed, plain):
    b = api.Batch(ctx, mixed)
    try:
        (tiles, rows), ms = b.autotune(api.default_params(rand_seed=1), reps=1)
        assert tiles in (2, 4, 8, 16, 32) and rows in (16, 32, 64, 128, 256) and ms > 0
        for k in range(len(mixed)):
            _same(b.result(k), plain[k])
    finally:
        b.free()
