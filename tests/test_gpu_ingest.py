"""GPU: ingest on the device (rbg_ingest_*) against the oracle's restatement
of the reference's parsers: records, names, lengths and bases of the FASTA,
the depth rows matched by name, the resident assembly's graph, the error
paths, the reference's own test data, and the 50,000-contig workload written
as FASTA and depth text, parsed on the device and built into a graph whose
digest must equal the committed one."""
import json
import os

import numpy as np
import pytest

from digests import record_of_result
from ingest_text import depth_text, fasta_text, names_for
from oracle import oracle as orc
from rabbitbin_amd import api, synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _oracle_inputs(text, mc, ms):
    recs, nseq = orc.fasta_parse(text, mc, ms)
    large = [(n, s) for n, c, s in recs if c == "large"]
    small = [(n, s) for n, c, s in recs if c == "small"]
    return recs, nseq, large, small


def _check_fasta(ig, text, mc, ms):
    recs, nseq, large, small = _oracle_inputs(text, mc, ms)
    inf = ig.info()
    assert inf["records"] == nseq
    assert inf["n_large"] == len(large) and inf["n_small"] == len(small)
    assert inf["names"] == [n for n, _ in large]
    assert inf["small_names"] == [n for n, _ in small]
    assert list(inf["lens"]) == [len(s) for _, s in large]
    assert list(inf["small_lens"]) == [len(s) for _, s in small]
    seqs, off, _, _ = ig.copy()
    assert seqs.tobytes() == b"".join(s for _, s in large)
    assert list(off) == list(np.cumsum([0] + [len(s) for _, s in large]))
    return large, small


def test_reference_fixture(ctx):
    fa = open(os.path.join(GOLD, "ref_contigs.fa"), "rb").read()
    dp = open(os.path.join(GOLD, "ref_contigs_depth.txt"), "rb").read()
    ig = api.Ingest(ctx, fa)
    try:
        large, _ = _check_fasta(ig, fa, 2500, 1000)
        ig.depth(dp)
        inf = ig.info()
        assert inf["S"] == 1 and inf["depth_rows"] == 1
        _, _, dep, _ = ig.copy()
        assert dep[0, 0].view(np.uint32) == 0x416373EB
    finally:
        ig.free()


def _cut(seqs, off, seed):
    """The contigs cut to lengths spread over the tiny, small and large
    classes (prefixes of 200 .. 6,000 bases)."""
    rng = np.random.default_rng(seed)
    n = len(off) - 1
    keep = np.minimum(np.diff(off).astype(np.int64), rng.integers(200, 6000, n))
    parts = [seqs[int(off[i]):int(off[i]) + int(keep[i])] for i in range(n)]
    return np.concatenate(parts), np.concatenate([[0], np.cumsum(keep)]).astype(np.uint64)


@pytest.mark.parametrize("seed", [1, 2])
def test_fasta_and_depth_equal_oracle(ctx, seed):
    n, S = 3000, 6
    seqs, off, depth, _ = synth.make_assembly(n, 20, np.log(2500), 1.1, S, seed=40 + seed)
    seqs, off = _cut(seqs, off, seed)
    names = names_for(n, b"NODE")
    text = fasta_text(seqs, off, names, seed=seed)
    lens = np.diff(off).astype(np.int64)
    dtext = depth_text(names, lens, depth, seed=seed, extra_rows=50, dup_rows=30)
    ig = api.Ingest(ctx, text, min_contig=2500, min_small_contig=1000)
    try:
        large, small = _check_fasta(ig, text, 2500, 1000)
        assert len(large) > 500 and len(small) > 300
        ig.depth(dtext)
        inf = ig.info()
        assert inf["S"] == S
        _, _, dep, sdep = ig.copy()
        want, found, rows = orc.depth_parse_match(dtext, [nm for nm, _ in large], S, min_depth_len=1000)
        assert found.all() and inf["depth_rows"] == rows
        assert np.array_equal(dep.view(np.uint32), want.view(np.uint32))
        swant, sfound, _ = orc.depth_parse_match(dtext, [nm for nm, _ in small], S, min_depth_len=1000)
        assert np.array_equal(sdep.view(np.uint32), swant.view(np.uint32))
        # the resident assembly from the ingest builds the graph of the same contigs and depths
        a = ig.assembly()
        try:
            g = a.build_graph(api.default_params(rand_seed=3))
        finally:
            a.free()
        lseqs = np.frombuffer(b"".join(s for _, s in large), np.uint8)
        loff = np.cumsum([0] + [len(s) for _, s in large]).astype(np.uint64)
        ref = ctx.build_graph(lseqs, loff, want, api.default_params(rand_seed=3))
        for k in ("from", "to", "scomp", "score"):
            assert np.array_equal(g[k], ref[k]), k
        assert g["sim_cutoff"] == ref["sim_cutoff"] and g["b0"] == ref["b0"]
    finally:
        ig.free()


def test_ingest_errors(ctx):
    seqs, off, depth, _ = synth.make_assembly(40, 5, np.log(4000), 0.3, 2, seed=9)
    names = names_for(40)
    text = fasta_text(seqs, off, names, seed=3)
    lens = np.diff(off).astype(np.int64)
    full = depth_text(names, lens, depth, seed=3)
    lines = full.split(b"\n")
    # a large contig without a row
    ig = api.Ingest(ctx, text, min_contig=100, min_small_contig=50)
    try:
        missing = b"\n".join(l for l in lines if not l.startswith(names[7] + b"\t"))
        with pytest.raises(api.RbgError):
            ig.depth(missing)
        # a hexadecimal mean in a kept row
        row = [l for l in lines if l.startswith(names[3] + b"\t")][0]
        f = row.split(b"\t")
        f[3] = b"0x1.8p1"
        with pytest.raises(api.RbgError):
            ig.depth(full.replace(row, b"\t".join(f)))
        ig.depth(full)  # the good file still parses afterwards
        assert ig.info()["S"] == 2
    finally:
        ig.free()
    # duplicate large-contig names
    dup = text + b">" + names[0] + b"\n" + b"A" * 200 + b"\n"
    ig = api.Ingest(ctx, dup, min_contig=100, min_small_contig=50)
    try:
        with pytest.raises(api.RbgError):
            ig.depth(full)
    finally:
        ig.free()


def test_asm50k_through_ingest_matches_committed_digest(ctx):
    """The 50,000-contig workload written as FASTA (wrapped at 0, 60 or 80
    columns, some CRLF) and as a depth file whose numbers print every float
    exactly (%.9g), parsed on the device with min_contig = 0 (every record
    large), gives the graph of the committed asm50k digest."""
    path = os.path.join(GOLD, "graph_digests.json")
    want = json.load(open(path))["asm50k"]
    seqs, off, depth, _ = synth.workload("asm50k")
    n = len(off) - 1
    names = names_for(n, b"k141")
    text = fasta_text(seqs, off, names, seed=7, widths=(0, 60, 80), crlf=0.05, blank=0.0)
    lens = np.diff(off).astype(np.int64)
    dtext = depth_text(names, lens, depth, seed=7, formats=("%.9g",))
    ig = api.Ingest(ctx, text, min_contig=0, min_small_contig=0)
    try:
        inf = ig.info()
        assert inf["n_large"] == n and inf["large_bases"] == int(off[-1])
        ig.depth(dtext, min_depth_len=0)
        inf = ig.info()
        print(f"asm50k ingest: FASTA {len(text) / 1e6:.0f} MB in {inf['ms_fasta']:.1f} ms, "
              f"depth {len(dtext) / 1e6:.0f} MB in {inf['ms_depth']:.1f} ms", flush=True)
        _, _, dep, _ = ig.copy()
        assert np.array_equal(dep.view(np.uint32), depth.view(np.uint32))
        a = ig.assembly()
        try:
            g = a.build_graph(api.default_params(rand_seed=1))
        finally:
            a.free()
    finally:
        ig.free()
    got = record_of_result(g)
    for k in ("n_edges", "sim_cutoff", "b0", "regime", "gated_pairs", "digest"):
        assert got[k] == want[k], k
