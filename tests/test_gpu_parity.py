"""GPU parity: librbg (through the C ABI) against the CPU oracle on the same
seeded inputs.  Integer/index outputs must be bit-exact; edge weights within
1e-6 of the FP64 oracle."""
import numpy as np
import pytest

from rabbitbin_amd import api, synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _edges(g, key="scomp"):
    return {(int(a), int(b)): float(s) for a, b, s in zip(g["from"], g["to"], g[key])}


def _check_graph(g, ref, near_tol=1e-6):
    assert np.array_equal(g["win16"], ref["win16"])
    assert g["b0"] == ref["b0"]
    assert np.array_equal(g["rowmax_cnt"], ref["rowmax_cnt"])
    assert np.array_equal(g["topk_len"], ref["topk_len"])
    n, K = g["topk_id"].shape
    for i in range(n):
        L = g["topk_len"][i]
        assert set(g["topk_id"][i, :L].tolist()) == set(ref["topk_id"][i, :L].tolist()), i
    assert g["sim_cutoff"] == ref["sim_cutoff"]
    eg = _edges(g)
    er = {(int(a), int(b)): float(s) for a, b, s in zip(ref["from"], ref["to"], ref["sv"])}
    assert eg == er
    _check_scores(g, ref, near_tol)


def _check_scores(g, ref, tol=1e-6):
    """Edge weights: bit-equal to the reference's arithmetic (score_ref); the
    near-cut list equal to the oracle's; off that list within tol of FP64,
    and on it the FP64 weights equal the oracle's."""
    if g["score"] is None:
        return
    assert np.array_equal(g["score"].view(np.uint32), ref["score_ref"].view(np.uint32))
    assert np.array_equal(g["near_edge"], ref["near"])
    off = np.ones(len(g["score"]), bool)
    off[ref["near"].astype(np.int64)] = False
    d = np.abs(g["score"].astype(np.float64) - ref["score64"])
    assert len(d) == 0 or d[off].max(initial=0.0) <= tol
    assert np.array_equal(g["near_score64"], ref["score64"][ref["near"].astype(np.int64)])


def test_sketch_winners_exact(ctx):
    seqs, off, _, _ = synth.make_assembly(300, 6, np.log(8000), 0.8, 0, seed=7)
    w32, w16 = ctx.build_pmh_winners(seqs, off, m=500, seed=42)
    ref32 = orc.sketch_batch(seqs, off, m=500, seed=42)
    assert np.array_equal(w32, ref32)
    assert np.array_equal(w16, orc.fold16(ref32))


def test_graph_small_regime_2k(ctx):
    seqs, off, depth, _ = synth.workload("asm2k")
    p = api.default_params(rand_seed=1)
    g = ctx.build_graph(seqs, off, depth, p, diag=True)
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=1)
    assert g["regime"] == 1
    _check_graph(g, ref)


@pytest.mark.parametrize("S", [0, 1, 2, 3, 10, 30, 40, 64])
def test_graph_fused_regime_forced_3k(ctx, S):
    # S selects the gate variant: 0, 1 and 2 -> no gate (composition weight
    # 1; every pair passing is_nz is counted into the dense matrix), 3..16
    # one 16-lane chunk, 30 two chunks in acc0, 40 a 32-wide chunk (acc1),
    # 64 two 32-wide chunks.
    seqs, off, depth, _ = synth.make_assembly(3000, 30, np.log(8000), 0.8, S, seed=99 + S)
    p = api.default_params(rand_seed=3, regime=2)
    g = ctx.build_graph(seqs, off, depth, p, diag=True)
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=3, force_regime=2)
    assert g["regime"] == 2
    assert g["gated_pairs"] == ref["gated_pairs"]
    _check_graph(g, ref)


@pytest.mark.parametrize("m,K,regime,S", [(100, 37, 2, 10), (256, 300, 2, 30), (512, 1024, 2, 16),
                                         (64, 5, 1, 3), (33, 1, 2, 3), (16, 8, 2, 10), (500, 199, 1, 0)])
def test_graph_sketch_sizes_and_max_edges(ctx, m, K, regime, S):
    """Other --sketch-m / --max-edges: unaligned K (segment loads of the mutual
    test), K > 256 (several keep-mask words, wider fence rows, a larger sort),
    K = 1, every plane-word count W in {2, 4, 8, 16} of the compare kernels."""
    seqs, off, depth, _ = synth.make_assembly(3000, 30, np.log(8000), 0.8, S, seed=1000 + m + K)
    p = api.default_params(rand_seed=4, regime=regime, m=m, max_edges=K)
    g = ctx.build_graph(seqs, off, depth, p, diag=True)
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=4, m=m, K=K, force_regime=regime)
    assert g["regime"] == regime and g["gated_pairs"] == ref["gated_pairs"]
    _check_graph(g, ref)


def _concat(contigs):
    off = np.zeros(len(contigs) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in contigs])
    return b"".join(contigs), off


EDGE_CONTIGS = [b"", b"A", b"ACG", b"ACGT", b"NNNN", b"acgtacgtAC", b"ACGTNACGTNNACGT", b"TTTTTTTTTTTTTTTT",
                b"ACGT" * 100, b"GGGGCCCC" * 50, b"AC-GT*RYac" * 20, bytes(range(256)) * 3, b"N" * 5000,
                b"ACGTTGCA" * 1000 + b"N" * 7 + b"ttgcaacg" * 500]


@pytest.mark.parametrize("m", [2, 16, 100, 512])
def test_sketch_edge_cases(ctx, m):
    """Empty, shorter than 4, all-N, lowercase, junk bytes, N runs, ragged
    lengths at every alignment of the sequence buffer."""
    rng = np.random.default_rng(m)
    contigs = list(EDGE_CONTIGS)
    for L in rng.integers(0, 3000, 40):
        s = np.array(list("ACGTacgtN"))[rng.choice(9, int(L), p=[.22, .22, .22, .22, .02, .02, .02, .02, .04])]
        contigs.append("".join(s).encode())
    seqs, off = _concat(contigs)
    w32, w16 = ctx.build_pmh_winners(seqs, off, m=m, seed=42)
    ref32 = orc.sketch_batch(seqs, off, m=m, seed=42)
    assert np.array_equal(w32, ref32)
    assert np.array_equal(w16, orc.fold16(ref32))
    assert not w32[:3].any() and not w32[4].any()  # no 4-mer: every register unset


def test_contig_length_limit(ctx):
    """The sketch addresses a workgroup's four contigs with 32-bit offsets:
    a contig of 512 MiB or more, or decreasing offsets, is refused before any
    byte is read or uploaded (the buffer here is 64 bytes)."""
    seqs = b"ACGT" * 16
    for off in ([0, 1 << 29], [0, 40, 20]):
        with pytest.raises(api.RbgError, match="512 MiB"):
            ctx.build_pmh_winners(seqs, np.array(off, dtype=np.uint64), m=500, seed=42)
    off = np.array([0, 64], dtype=np.uint64)
    w32, _ = ctx.build_pmh_winners(seqs, off, m=500, seed=42)
    assert np.array_equal(w32, orc.sketch_batch(seqs, off, m=500, seed=42))


def test_graph_natural_fused_regime_26k(ctx):
    """n = 26,000 > 25,000 selects the fused regime by itself (S = 10: one
    masked 16-lane chunk in the gate); the resident assembly with the edge
    list kept in HBM gives the same graph."""
    seqs, off, depth, _ = synth.workload("asm50k", n=26000)
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=1)
    assert ref["regime"] == 2
    g = ctx.build_graph(seqs, off, depth, api.default_params(rand_seed=1), diag=True)
    assert g["regime"] == 2 and g["gated_pairs"] == ref["gated_pairs"]
    _check_graph(g, ref)
    asm = api.Assembly(ctx, seqs, off, depth)
    try:
        r = asm.build_graph(api.default_params(rand_seed=1), on_device=True)
        assert r["from"] is None and r["n_edges"] == len(g["from"])
        frm, to, sc, score = asm.copy_edges(r["n_edges"])
        assert np.array_equal(frm, g["from"]) and np.array_equal(to, g["to"])
        assert np.array_equal(sc, g["scomp"]) and np.array_equal(score, g["score"])
        assert np.array_equal(r["near_edge"], g["near_edge"])
    finally:
        asm.free()


def test_non_default_semantics_refused(ctx, monkeypatch):
    """A RABBIT_* variable that changes the path's results makes the build
    refuse (the caller routes the run to the reference's CPU path)."""
    seqs, off, depth, _ = synth.make_assembly(600, 12, np.log(8000), 0.8, 3, seed=5)
    monkeypatch.setenv("RABBIT_MUTUAL_KNN", "0")
    p = api.default_params(rand_seed=1)
    assert p.semantics == api.RBG_SEM_NO_MUTUAL
    with pytest.raises(api.RbgError, match="RABBIT_MUTUAL_KNN"):
        ctx.build_graph(seqs, off, depth, p)
    monkeypatch.delenv("RABBIT_MUTUAL_KNN")
    g = ctx.build_graph(seqs, off, depth, api.default_params(rand_seed=1))
    assert g["n_edges"] > 0


@pytest.mark.parametrize("S", [16, 30])
def test_near_cut_list(ctx, S):
    """With a wide tolerance the near-cut list is long: it must equal the
    oracle's, and every edge off it must agree with FP64 to that tolerance."""
    seqs, off, depth, _ = synth.make_assembly(3000, 30, np.log(8000), 0.8, S, seed=99 + S)
    # depths with multiplicative noise spread the weights over [0, 1] (the
    # clean synthetic depths put nearly every weight above 0.9)
    depth = (depth * np.random.default_rng(S).lognormal(0, 0.6, depth.shape)).astype(np.float32)
    tol = 0.05
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=3, force_regime=2, near_tol=tol)
    g = ctx.build_graph(seqs, off, depth, api.default_params(rand_seed=3, regime=2, near_tol=tol), diag=True)
    assert len(ref["near"]) > 1000
    _check_graph(g, ref, near_tol=tol)


def test_edge_weights_entry_point(ctx):
    """rbg_edge_weights on an arbitrary edge list: score bit-equal to the
    reference's arithmetic and score64 to the FP64 oracle, for S on both sides
    of the 16- and 32-wide chunking."""
    rng = np.random.default_rng(5)
    for S in (2, 3, 16, 30, 32, 47, 64, 70, 100):
        n = 500
        d = rng.poisson(rng.lognormal(1, 1, (n, 1)) * rng.lognormal(0, 1, (1, S)) * 4).astype(np.float32)
        _, unit, raw, colnorm, _ = orc.depth_features(d)
        ne = 5000
        fr = rng.integers(0, n, ne).astype(np.uint32)
        to = rng.integers(0, n, ne).astype(np.uint32)
        sc = rng.random(ne).astype(np.float32)
        w, w64 = ctx.edge_weights(fr, to, sc, unit, raw, colnorm, S, orc.w_comp_for(S))
        _, rw64 = orc.edge_weights(fr, to, sc, unit, raw, colnorm, S, mode=1)
        rw, _ = orc.edge_weights(fr, to, sc, unit, raw, colnorm, S, mode=0)
        assert np.array_equal(w64, rw64) and np.array_equal(w.view(np.uint32), rw.view(np.uint32))


def test_full_size_assembly_properties(ctx):
    """asm300k (300,000 contigs, S = 30): size-independent checks of the whole
    build -- sampled sketches, b0 and calibration recomputed from the GPU's
    own outputs, exact top-K / row maxima of sampled rows against all
    partners, and for sampled edges the count, the gate and the weight."""
    seqs, off, depth, _ = synth.workload("asm300k")
    n = len(off) - 1
    asm = api.Assembly(ctx, seqs, off, depth)
    try:
        g = asm.build_graph(api.default_params(rand_seed=1), diag=True)
    finally:
        asm.free()
    frm, to, sc, score = g["from"], g["to"], g["scomp"], g["score"]
    assert g["regime"] == 2 and len(frm) > 1_000_000
    # edges: sorted, i < j, unique, above the cutoff
    assert (frm < to).all()
    key = frm.astype(np.uint64) << np.uint64(32) | to
    assert (np.diff(key.astype(np.int64)) > 0).all()
    assert (sc.astype(np.float64) > g["sim_cutoff"] / 1000.0).all()
    # sketches of sampled contigs
    rng = np.random.default_rng(300)
    sample = np.sort(rng.choice(n, 400, replace=False))
    buf = np.asarray(seqs, np.uint8) if not isinstance(seqs, bytes) else np.frombuffer(seqs, np.uint8)
    sub = [bytes(buf[off[c]:off[c + 1]]) for c in sample]
    s_seqs, s_off = _concat(sub)
    assert np.array_equal(g["win16"][sample], orc.fold16(orc.sketch_batch(s_seqs, s_off)))
    # b0 and calibration from the GPU's own winners / row maxima
    assert orc.pmh_baseline(g["win16"]) == g["b0"]
    assert orc.calibrate(g["rowmax_cnt"], 500, g["b0"], 2, 0.95, 1) == g["sim_cutoff"]
    # exact top-K, row maxima and gated-partner counts of sampled rows
    _, unit, raw, colnorm, anynz = orc.depth_features(depth)
    rows = np.sort(rng.choice(n, 48, replace=False)).astype(np.uint32)
    ids, ln, rm, _ = orc.rows_topk(g["win16"], unit, anynz, 30, g["b0"], rows)
    for t, r in enumerate(rows):
        assert g["topk_len"][r] == ln[t] and g["rowmax_cnt"][r] == rm[t]
        assert np.array_equal(np.sort(g["topk_id"][r, :ln[t]]), ids[t, :ln[t]])
    # mutual filter, counts, gate and weights of sampled edges
    es = rng.choice(len(frm), 3000, replace=False)
    for e in es[:500]:
        i, j = int(frm[e]), int(to[e])
        assert j in g["topk_id"][i, :g["topk_len"][i]] and i in g["topk_id"][j, :g["topk_len"][j]]
    svd = np.array([orc.sv_from_count(c, 500, g["b0"]) for c in range(501)])
    for e in es:
        i, j = int(frm[e]), int(to[e])
        c = orc.orc().orc_match_count16(g["win16"][i].ctypes.data_as(orc.u16p),
                                        g["win16"][j].ctypes.data_as(orc.u16p), 500)
        assert np.float32(svd[c]) == sc[e]
        assert not (orc.unit_dot_f(unit[i], unit[j]) < np.float32(0.7 - 1e-4))
    w, w64 = orc.edge_weights(frm[es], to[es], sc[es], unit, raw, colnorm, 30, mode=1)
    wr, _ = orc.edge_weights(frm[es], to[es], sc[es], unit, raw, colnorm, 30, mode=0)
    assert np.array_equal(score[es].view(np.uint32), wr.view(np.uint32))
    near = set(g["near_edge"].tolist())
    keep = np.array([e not in near for e in es])
    assert np.abs(score[es].astype(np.float64) - w64)[keep].max() <= 1e-6


@pytest.mark.parametrize("S,regime", [(3, 1), (10, 2), (30, 2), (0, 1)])
def test_graph_from_sketches(ctx, S, regime):
    """rbg_build_graph_from_sketches fed the oracle's g_win16, depth rows and
    anynz gives the oracle's graph (b0 estimated, and b0 given)."""
    seqs, off, depth, _ = synth.make_assembly(3000, 30, np.log(8000), 0.8, S, seed=500 + S)
    ref = orc.full_pipeline(seqs, off, depth, rand_seed=2, force_regime=regime)
    kw = {}
    if S:
        _, unit, raw, colnorm, anynz = orc.depth_features(depth)
        kw = dict(unit=unit, raw=raw, colnorm=colnorm if S > 1 else None, anynz=anynz)
    p = api.default_params(rand_seed=2, regime=regime)
    g = ctx.build_graph_from_sketches(ref["win16"], params=p, diag=True, **kw)
    g["win16"] = ref["win16"]
    assert g["regime"] == regime
    _check_graph(g, ref)
    g2 = ctx.build_graph_from_sketches(ref["win16"], b0=ref["b0"], params=p, **kw)
    assert g2["b0"] == ref["b0"] and np.array_equal(g2["from"], g["from"]) and np.array_equal(g2["to"], g["to"])


def test_graph_from_arbitrary_winners(ctx):
    """Winner values that are not this library's own fold (random 16-bit
    relabelling per register) give the same graph: the class numbering only
    keeps equality."""
    seqs, off, depth, _ = synth.make_assembly(2600, 25, np.log(8000), 0.8, 10, seed=77)
    w32 = orc.sketch_batch(seqs, off)
    w16 = orc.fold16(w32)
    rng = np.random.default_rng(1)
    relabel = rng.permutation(65535).astype(np.uint16) + 1  # non-zero -> non-zero, injective
    wr = np.where(w16 == 0, 0, relabel[np.maximum(w16.astype(np.int64) - 1, 0)]).astype(np.uint16)
    _, unit, raw, colnorm, anynz = orc.depth_features(depth)
    b0 = orc.pmh_baseline(wr)
    ref = orc.build_graph(wr, unit, anynz, 10, b0, rand_seed=1, force_regime=2)
    g = ctx.build_graph_from_sketches(wr, unit=unit, raw=raw, colnorm=colnorm, anynz=anynz,
                                      params=api.default_params(rand_seed=1, regime=2))
    assert g["b0"] == b0
    assert set(zip(g["from"].tolist(), g["to"].tolist())) == set(zip(ref["from"].tolist(), ref["to"].tolist()))


def test_graph_round_trips_through_cache(ctx, tmp_path):
    """A graph built on the GPU written as an RBCACHE1 v3 file (with the
    ranked depth matrix, raw depths and column norms the reference caches)
    reads back unchanged."""
    seqs, off, depth, _ = synth.make_assembly(2600, 25, np.log(8000), 0.8, 10, seed=31)
    g = ctx.build_graph(seqs, off, depth, api.default_params(rand_seed=1, regime=2))
    ranked, _, raw, colnorm, _ = api.depth_features(depth)
    n = len(off) - 1
    p = str(tmp_path / "asm.cache")
    api.write_cache(p, nobs=n, nobs1=0, samples=10, seed=1, total_size=int(off[-1]), total_size1=0,
                    sim_cutoff=g["sim_cutoff"], b0=g["b0"], min_contig=2500, min_small_contig=1000, has_depth=1,
                    names=[f"c{i}" for i in range(n)], small_names=[], seq_lens=np.diff(off.astype(np.int64)),
                    small_seq_lens=np.zeros(0, np.uint64), depth=ranked, small_depth=np.zeros((0, 10), np.float32),
                    large_means=depth.ravel(), depth_raw=raw.ravel(), colnorm=colnorm, from_=g["from"], to=g["to"],
                    scomp=g["scomp"])
    c = api.read_cache(p)
    assert c["sim_cutoff"] == g["sim_cutoff"] and c["b0"] == g["b0"] and c["n_edges"] == g["n_edges"]
    assert np.array_equal(c["from_"], g["from"]) and np.array_equal(c["to"], g["to"])
    assert np.array_equal(c["scomp"], g["scomp"]) and np.array_equal(c["depth"], ranked)
