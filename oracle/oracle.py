"""Python binding of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_probmh.so")
PORT512_LIB = os.path.join(HERE, "liboracle_port512.so")

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class GraphIn(C.Structure):
    _fields_ = [("win16", u16p), ("unit", f32p), ("anynz", u8p), ("n", C.c_uint64), ("m", C.c_uint32),
                ("S", C.c_uint32), ("K", C.c_uint32), ("min_edge_weight", C.c_double),
                ("coverage", C.c_double), ("w_comp", C.c_double), ("b0", C.c_double),
                ("rand_seed", C.c_uint), ("threads", C.c_int), ("force_regime", C.c_int)]


class GraphOut(C.Structure):
    _fields_ = [("sim_cutoff", C.c_double), ("n_edges", C.c_uint64), ("from_", u32p), ("to", u32p),
                ("sv", f32p), ("topk_id", u32p), ("topk_sv", f32p), ("topk_len", u32p),
                ("rowmax_cnt", u32p), ("regime", C.c_int), ("gated_pairs", C.c_uint64)]


class PortStats(C.Structure):
    _fields_ = [("prep_s", C.c_double), ("pass_s", C.c_double), ("calib_s", C.c_double),
                ("mutual_s", C.c_double), ("tile", C.c_uint64), ("tilepairs", C.c_uint64),
                ("q8_pruned", C.c_uint64), ("float_checked", C.c_uint64), ("ub_pruned", C.c_uint64),
                ("compares", C.c_uint64), ("avx512", C.c_int), ("q8", C.c_int)]


_orc = None
_ref = None
_port512 = None


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def orc():
    global _orc
    if _orc is None:
        L = C.CDLL(ORACLE_LIB)
        vp = C.c_void_p
        L.orc_fmix64.restype = C.c_uint64
        L.orc_fmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.orc_wyrand_seq.argtypes = [u64p, C.c_int, u64p]
        L.orc_wy_uniform_int_seq.argtypes = [u64p, C.c_uint32, C.c_int, u32p]
        L.orc_zig_seq.argtypes = [u64p, C.c_int, f64p]
        L.orc_ted_seq.argtypes = [C.c_double, C.c_double, C.c_double, u64p, C.c_int, f64p]
        L.orc_ted_params.argtypes = [C.c_uint32, f64p, f64p]
        L.orc_pmh_new.restype = vp
        L.orc_pmh_new.argtypes = [C.c_uint32, C.c_uint64]
        L.orc_pmh_free.argtypes = [vp]
        L.orc_pmh_reset.argtypes = [vp]
        L.orc_pmh_add_hash.argtypes = [vp, C.c_uint64, C.c_double]
        L.orc_pmh_get.argtypes = [vp, u64p, f64p]
        L.orc_kmer4_hist.argtypes = [C.c_char_p, C.c_size_t, u32p, u64p, u64p]
        L.orc_kmer4_weights.argtypes = [u32p, u64p, C.c_uint64, C.c_double, f64p]
        L.orc_build_pmh_winners.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.c_uint64, u32p]
        L.orc_sketch_batch.argtypes = [C.c_char_p, u64p, C.c_size_t, C.c_uint32, C.c_uint64, u32p, C.c_int]
        L.orc_fold16.argtypes = [u32p, C.c_size_t, u16p]
        L.orc_depth_features.argtypes = [f32p, C.c_size_t, C.c_size_t, C.c_double, f32p, f32p, f32p, f64p, u8p]
        L.orc_match_count16.restype = C.c_uint32
        L.orc_match_count16.argtypes = [u16p, u16p, C.c_uint32]
        L.orc_pmh_baseline.restype = C.c_double
        L.orc_pmh_baseline.argtypes = [u16p, C.c_size_t, C.c_uint32]
        L.orc_unit_dot_f.restype = C.c_float
        L.orc_unit_dot_f.argtypes = [f32p, f32p, C.c_uint32]
        L.orc_sv_from_count.restype = C.c_double
        L.orc_sv_from_count.argtypes = [C.c_uint32, C.c_uint32, C.c_double]
        L.orc_rand_seq.argtypes = [C.c_uint, C.c_int, C.POINTER(C.c_int32)]
        L.orc_build_graph.argtypes = [C.POINTER(GraphIn), C.POINTER(GraphOut)]
        L.orc_graph_free.argtypes = [C.POINTER(GraphOut)]
        L.orc_edge_weights.argtypes = [u32p, u32p, f32p, C.c_uint64, f32p, f32p, f64p, C.c_uint32, C.c_double,
                                       C.c_double, C.c_double, C.c_int, f32p, f64p]
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB)
        vp = C.c_void_p
        L.ref_pmh_new.restype = vp
        L.ref_pmh_new.argtypes = [C.c_uint32, C.c_int, C.c_uint64]
        L.ref_pmh_free.argtypes = [vp]
        L.ref_pmh_reset.argtypes = [vp]
        L.ref_pmh_add_hash.argtypes = [vp, C.c_uint64, C.c_double]
        L.ref_pmh_get.argtypes = [vp, u64p, f64p]
        L.ref_wyrand_seq.argtypes = [u64p, C.c_int, u64p]
        L.ref_wy_uniform_int_seq.argtypes = [u64p, C.c_uint32, C.c_int, u32p]
        L.ref_zig_seq.argtypes = [u64p, C.c_int, f64p]
        L.ref_ted_seq.argtypes = [C.c_double, C.c_double, C.c_double, u64p, C.c_int, f64p]
        L.ref_fmix64.restype = C.c_uint64
        L.ref_fmix64.argtypes = [C.c_uint64, C.c_uint64]
        L.ref_zig_tables.argtypes = [f64p, f64p]
        _ref = L
    return _ref


# ---- convenience wrappers -------------------------------------------------

def sketch_batch(seqs, offsets, m=500, seed=42, threads=0):
    buf = bytes(np.ascontiguousarray(seqs, np.uint8)) if not isinstance(seqs, bytes) else seqs
    off = np.ascontiguousarray(offsets, np.uint64)
    n = len(off) - 1
    out = np.zeros((n, m), np.uint32)
    orc().orc_sketch_batch(buf, _p(off, u64p), n, m, seed, _p(out, u32p), threads)
    return out


def fold16(w32):
    w = np.ascontiguousarray(w32, np.uint32)
    out = np.zeros(w.shape, np.uint16)
    orc().orc_fold16(_p(w, u32p), w.size, _p(out, u16p))
    return out


def depth_features(depth, min_cv=1.0):
    d = np.ascontiguousarray(depth, np.float32)
    n, S = d.shape
    ranked = np.zeros_like(d)
    unit = np.zeros_like(d)
    raw = np.zeros_like(d)
    colnorm = np.zeros(S, np.float64)
    anynz = np.zeros(n, np.uint8)
    orc().orc_depth_features(_p(d, f32p), n, S, min_cv, _p(ranked, f32p), _p(unit, f32p), _p(raw, f32p),
                             _p(colnorm, f64p), _p(anynz, u8p))
    return ranked, unit, raw, colnorm, anynz


def pmh_baseline(win16):
    w = np.ascontiguousarray(win16, np.uint16)
    return orc().orc_pmh_baseline(_p(w, u16p), w.shape[0], w.shape[1])


def w_comp_for(S):
    # auto_w_comp (src/rabbitbin.cpp:470-473)
    return 1.0 if S <= 2 else 0.0


def build_graph(win16, unit, anynz, S, b0, rand_seed=1, K=200, min_edge_weight=0.70, coverage=0.95,
                force_regime=0, threads=0):
    w = np.ascontiguousarray(win16, np.uint16)
    n, m = w.shape
    u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
    nz = np.ascontiguousarray(anynz, np.uint8) if anynz is not None else None
    gi = GraphIn(_p(w, u16p), _p(u, f32p), _p(nz, u8p), n, m, S, K, min_edge_weight, coverage,
                 w_comp_for(S), b0, rand_seed, threads, force_regime)
    go = GraphOut()
    st = orc().orc_build_graph(C.byref(gi), C.byref(go))
    if st != 0:
        raise RuntimeError(f"orc_build_graph status {st}")
    ne = go.n_edges
    res = {
        "sim_cutoff": go.sim_cutoff, "regime": go.regime, "gated_pairs": go.gated_pairs,
        "from": np.ctypeslib.as_array(go.from_, (ne,)).copy() if ne else np.zeros(0, np.uint32),
        "to": np.ctypeslib.as_array(go.to, (ne,)).copy() if ne else np.zeros(0, np.uint32),
        "sv": np.ctypeslib.as_array(go.sv, (ne,)).copy() if ne else np.zeros(0, np.float32),
        "topk_id": np.ctypeslib.as_array(go.topk_id, (n, K)).copy(),
        "topk_sv": np.ctypeslib.as_array(go.topk_sv, (n, K)).copy(),
        "topk_len": np.ctypeslib.as_array(go.topk_len, (n,)).copy(),
        "rowmax_cnt": np.ctypeslib.as_array(go.rowmax_cnt, (n,)).copy(),
    }
    orc().orc_graph_free(C.byref(go))
    return res


def edge_weights(frm, to, scomp, unit, raw, colnorm, S, mode, min_edge_weight=0.70, neg_thr=-0.3):
    frm = np.ascontiguousarray(frm, np.uint32)
    to = np.ascontiguousarray(to, np.uint32)
    sc = np.ascontiguousarray(scomp, np.float32)
    u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
    r = np.ascontiguousarray(raw, np.float32) if raw is not None else None
    cn = np.ascontiguousarray(colnorm, np.float64) if colnorm is not None else None
    w = np.zeros(len(frm), np.float32)
    w64 = np.zeros(len(frm), np.float64)
    orc().orc_edge_weights(_p(frm, u32p), _p(to, u32p), _p(sc, f32p), len(frm), _p(u, f32p), _p(r, f32p),
                           _p(cn, f64p), S, w_comp_for(S), min_edge_weight, neg_thr, mode, _p(w, f32p),
                           _p(w64, f64p))
    return w, w64


def edge_near(frm, to, scomp, unit, raw, colnorm, S, tol, min_edge_weight=0.70, neg_thr=-0.3):
    """Indices of the edges whose FP64 correlation is within tol of the
    negative cut or whose FP64 weight before the cut is within tol of the
    min-edge cut (ascending)."""
    frm = np.ascontiguousarray(frm, np.uint32)
    to = np.ascontiguousarray(to, np.uint32)
    sc = np.ascontiguousarray(scomp, np.float32)
    u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
    r = np.ascontiguousarray(raw, np.float32) if raw is not None else None
    cn = np.ascontiguousarray(colnorm, np.float64) if colnorm is not None else None
    near = np.zeros(len(frm), np.uint8)
    L = orc()
    L.orc_edge_near.argtypes = [u32p, u32p, f32p, C.c_uint64, f32p, f32p, f64p, C.c_uint32, C.c_double,
                                C.c_double, C.c_double, C.c_double, u8p]
    L.orc_edge_near(_p(frm, u32p), _p(to, u32p), _p(sc, f32p), len(frm), _p(u, f32p), _p(r, f32p),
                    _p(cn, f64p), S, w_comp_for(S), min_edge_weight, neg_thr, tol, _p(near, u8p))
    return np.flatnonzero(near).astype(np.uint64)


def chi2_cdf(m, x):
    """CDF of the chi-squared distribution with 2m degrees of freedom at x, in
    the reference's Poisson-sum form."""
    L = orc()
    L.orc_chi2_cdf.restype = C.c_double
    L.orc_chi2_cdf.argtypes = [C.c_int, C.c_double]
    return L.orc_chi2_cdf(m, x)


def lpa_node_order(n, seed):
    L = orc()
    L.orc_lpa_node_order.argtypes = [C.c_uint64, C.c_uint64, u64p]
    out = np.zeros(n, np.uint64)
    L.orc_lpa_node_order(n, seed, _p(out, u64p))
    return out


def label_propagation(frm, to, score, n, seed, with_frozen=False):
    """Fisher label propagation (the reference's default sequential sweep):
    (membership, rounds[, frozen flags])."""
    L = orc()
    L.orc_label_propagation.restype = C.c_int
    L.orc_label_propagation.argtypes = [u32p, u32p, f32p, C.c_uint64, C.c_uint64, C.c_uint64, u64p, u8p]
    frm = np.ascontiguousarray(frm, np.uint32)
    to = np.ascontiguousarray(to, np.uint32)
    sc = np.ascontiguousarray(score, np.float32)
    out = np.zeros(n, np.uint64)
    frozen = np.zeros(n, np.uint8)
    r = L.orc_label_propagation(_p(frm, u32p), _p(to, u32p), _p(sc, f32p), len(frm), n, seed, _p(out, u64p),
                                _p(frozen, u8p))
    return (out, r, frozen) if with_frozen else (out, r)


class FaRec(C.Structure):
    _fields_ = [("name_beg", C.c_uint64), ("seq_off", C.c_uint64), ("seq_len", C.c_uint64),
                ("name_len", C.c_uint32), ("cls", C.c_uint32)]


def _bins_csr(bins):
    ids = np.array(list(bins.keys()), dtype=np.int32)
    sizes = [len(v) for v in bins.values()]
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes, dtype=np.uint64)
    mem = np.concatenate([np.asarray(v, np.uint64) for v in bins.values()]) if sizes else np.zeros(0, np.uint64)
    return ids, off, np.ascontiguousarray(mem, np.uint64)


def recruit_cores(S, unit_large, len_large, len_small, bins, core_floor, min_cs=10):
    """The frozen cores: (ids int32, thresholds f64, centroids f64 [cores x S])."""
    L = orc()
    i32p = C.POINTER(C.c_int32)
    L.orc_recruit_cores.restype = C.c_uint64
    L.orc_recruit_cores.argtypes = [C.c_uint32, C.c_uint32, f32p, u64p, u64p, C.c_uint32, i32p, u64p, u64p,
                                    C.c_uint64, C.c_uint64, C.c_uint64, i32p, f64p, f64p]
    ul = np.ascontiguousarray(unit_large, np.float32).reshape(-1)
    ll = np.ascontiguousarray(len_large, np.uint64)
    ls = np.ascontiguousarray(len_small, np.uint64)
    ids, off, mem = _bins_csr(bins)
    cap = max(1, len(ids))
    oid = np.zeros(cap, np.int32)
    thr = np.zeros(cap, np.float64)
    cent = np.zeros(cap * S, np.float64)
    n = L.orc_recruit_cores(S, len(ll), _p(ul, f32p), _p(ll, u64p), _p(ls, u64p), len(ids), _p(ids, i32p),
                            _p(off, u64p), _p(mem, u64p), core_floor, min_cs, cap, _p(oid, i32p), _p(thr, f64p),
                            _p(cent, f64p))
    return oid[:n], thr[:n], cent[:n * S].reshape(n, S)


def recruit_scores(S, unit, cent):
    """u . centroid of every (row, core) in the reference's order (n x cores)."""
    L = orc()
    L.orc_recruit_scores.argtypes = [C.c_uint32, C.c_uint64, f32p, C.c_uint64, f64p, f64p]
    u = np.ascontiguousarray(unit, np.float32).reshape(-1, S)
    c = np.ascontiguousarray(cent, np.float64).reshape(-1, S)
    out = np.zeros((u.shape[0], c.shape[0]), np.float64)
    L.orc_recruit_scores(S, u.shape[0], _p(u.reshape(-1), f32p), c.shape[0], _p(c.reshape(-1), f64p),
                         _p(out.reshape(-1), f64p))
    return out


def recruit_frozen_cores(S, unit_large, len_large, depth_small, len_small, bins, core_floor, min_cs=10):
    """(assign_large, assign_small, stats dict) of the reference's frozen-core recruitment."""
    L = orc()
    i32p = C.POINTER(C.c_int32)
    L.orc_recruit_frozen_cores.argtypes = [C.c_uint32, C.c_uint32, f32p, u64p, C.c_uint32, f32p, u64p,
                                           C.c_uint32, i32p, u64p, u64p, C.c_uint64, C.c_uint64, i32p, i32p, u64p]
    ul = np.ascontiguousarray(unit_large, np.float32).reshape(-1)
    ll = np.ascontiguousarray(len_large, np.uint64)
    ds = np.ascontiguousarray(depth_small, np.float32).reshape(-1)
    ls = np.ascontiguousarray(len_small, np.uint64)
    ids, off, mem = _bins_csr(bins)
    al = np.zeros(len(ll), np.int32)
    as_ = np.zeros(len(ls), np.int32)
    st = np.zeros(7, np.uint64)
    L.orc_recruit_frozen_cores(S, len(ll), _p(ul, f32p), _p(ll, u64p), len(ls), _p(ds, f32p), _p(ls, u64p),
                               len(ids), _p(ids, i32p), _p(off, u64p), _p(mem, u64p), core_floor, min_cs,
                               _p(al, i32p), _p(as_, i32p), _p(st, u64p))
    names = ("cores", "long_assigned", "long_ambiguous", "long_unmatched", "short_assigned", "short_ambiguous",
             "short_unmatched")
    return al, as_, {k: int(v) for k, v in zip(names, st)}


def fasta_parse(data, min_contig=2500, min_small=1000):
    """The reference's FASTA parse of `data` (bytes): a list of records
    (name, class 'tiny' | 'small' | 'large', sequence bytes or None for tiny)
    in file order, and num_seqs."""
    L = orc()
    L.orc_fasta_parse.restype = C.c_uint64
    L.orc_fasta_parse.argtypes = [C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(FaRec), C.c_uint64,
                                  C.c_char_p, u64p]
    buf = bytes(data)
    cap = buf.count(b">") + 1
    recs = (FaRec * cap)()
    bases = C.create_string_buffer(len(buf) + 1)
    ns = np.zeros(1, np.uint64)
    nr = L.orc_fasta_parse(buf, len(buf), min_contig, min_small, recs, cap, bases, _p(ns, u64p))
    out = []
    raw = bases.raw
    for k in range(nr):
        r = recs[k]
        name = buf[r.name_beg:r.name_beg + r.name_len]
        cls = {1: "tiny", 2: "small", 3: "large"}[r.cls]
        seq = raw[r.seq_off:r.seq_off + r.seq_len] if r.cls > 1 else None
        out.append((name, cls, seq))
    return out, int(ns[0])


def depth_parse_match(data, names, S, cv_ext=False, min_depth_len=1000):
    """The reference's depth-file parse (strtod / strtof) and the row of each
    name: (n x S float32 means, found flags, rows in the map)."""
    L = orc()
    L.orc_depth_parse_match.restype = C.c_uint64
    L.orc_depth_parse_match.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.c_int, C.c_uint64, C.c_char_p, u64p,
                                        C.c_uint64, f32p, u8p]
    buf = bytes(data)
    blob = b"".join(names)
    off = np.zeros(len(names) + 1, np.uint64)
    np.cumsum([len(x) for x in names], out=off[1:])
    out = np.zeros((len(names), S), np.float32)
    found = np.zeros(len(names), np.uint8)
    rows = L.orc_depth_parse_match(buf, len(buf), S, int(cv_ext), min_depth_len, blob, _p(off, u64p), len(names),
                                   _p(out, f32p), _p(found, u8p))
    return out, found.astype(bool), int(rows)


def depth_ncols(data):
    L = orc()
    L.orc_depth_ncols.restype = C.c_uint64
    L.orc_depth_ncols.argtypes = [C.c_char_p, C.c_uint64]
    buf = bytes(data)
    return int(L.orc_depth_ncols(buf, len(buf)))


def host_avx512_vnni():
    """True when the host CPU has AVX-512 F/BW/VNNI (from /proc/cpuinfo)."""
    try:
        flags = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
        return {"avx512f", "avx512bw", "avx512_vnni"} <= flags
    except OSError:
        return False


def _port_fn(avx512):
    global _port512
    L = orc()
    if avx512:
        if _port512 is None:
            _port512 = C.CDLL(PORT512_LIB)
            _port512.orc_port_fused_graph_avx512.argtypes = [C.POINTER(GraphIn), C.POINTER(GraphOut),
                                                             C.POINTER(PortStats)]
        return _port512.orc_port_fused_graph_avx512
    L.orc_port_fused_graph.argtypes = [C.POINTER(GraphIn), C.POINTER(GraphOut), C.POINTER(PortStats)]
    return L.orc_port_fused_graph


def port_fused_graph(win16, unit, anynz, S, b0, rand_seed=1, K=200, min_edge_weight=0.70, coverage=0.95,
                     threads=0, avx512=None):
    """The fused regime by the performance port of the reference's fused pass
    (triangle tile pass, int8 prefilter, composition early exit, spin-locked
    row heaps, mutual drain).  avx512=None uses the AVX-512 build where the
    host has it.  Edges come back in the port's emission order."""
    if avx512 is None:
        avx512 = host_avx512_vnni() and os.path.exists(PORT512_LIB)
    w = np.ascontiguousarray(win16, np.uint16)
    n, m = w.shape
    u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
    nz = np.ascontiguousarray(anynz, np.uint8) if anynz is not None else None
    gi = GraphIn(_p(w, u16p), _p(u, f32p), _p(nz, u8p), n, m, S, K, min_edge_weight, coverage,
                 w_comp_for(S), b0, rand_seed, threads, 2)
    go = GraphOut()
    ps = PortStats()
    st = _port_fn(avx512)(C.byref(gi), C.byref(go), C.byref(ps))
    if st != 0:
        raise RuntimeError(f"orc_port_fused_graph status {st}")
    ne = go.n_edges
    res = {
        "sim_cutoff": go.sim_cutoff, "regime": go.regime, "gated_pairs": go.gated_pairs,
        "from": np.ctypeslib.as_array(go.from_, (ne,)).copy() if ne else np.zeros(0, np.uint32),
        "to": np.ctypeslib.as_array(go.to, (ne,)).copy() if ne else np.zeros(0, np.uint32),
        "sv": np.ctypeslib.as_array(go.sv, (ne,)).copy() if ne else np.zeros(0, np.float32),
        "stats": {k: getattr(ps, k) for k, _ in PortStats._fields_},
    }
    orc().orc_graph_free(C.byref(go))
    return res


def omp_info():
    """OpenMP placement of liboracle.so: (proc_bind, places, max_threads)."""
    L = orc()
    L.orc_omp_info.argtypes = [C.POINTER(C.c_int)] * 3
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    L.orc_omp_info(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def pair_pass_rows(win16, unit, anynz, S, b0, row0, row1, K=200, threads=0):
    """Timed fused pair pass of rows [row0, row1) against all partners."""
    L = orc()
    L.orc_pair_pass_rows.restype = C.c_double
    L.orc_pair_pass_rows.argtypes = [u16p, f32p, u8p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                     C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, u64p, u64p]
    w = np.ascontiguousarray(win16, np.uint16)
    n, m = w.shape
    u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
    nz = np.ascontiguousarray(anynz, np.uint8) if anynz is not None else None
    pairs = C.c_uint64()
    gated = C.c_uint64()
    secs = L.orc_pair_pass_rows(_p(w, u16p), _p(u, f32p), _p(nz, u8p), n, m, S, b0, K, row0, row1, threads,
                                C.byref(pairs), C.byref(gated))
    return {"seconds": secs, "pairs": pairs.value, "gated": gated.value}


def full_pipeline(seqs, offsets, depth, rand_seed=1, m=500, K=200, force_regime=0, threads=0, near_tol=1e-6):
    """Oracle of the whole path: sketches, b0, graph, edge weights (reference arithmetic)."""
    w32 = sketch_batch(seqs, offsets, m=m, threads=threads)
    w16 = fold16(w32)
    b0 = pmh_baseline(w16)
    S = 0 if depth is None else depth.shape[1]
    if S:
        _, unit, raw, colnorm, anynz = depth_features(depth)
    else:
        unit = raw = colnorm = anynz = None
    g = build_graph(w16, unit, anynz, S, b0, rand_seed=rand_seed, K=K, force_regime=force_regime,
                    threads=threads)
    g["b0"] = b0
    g["win16"] = w16
    if S:
        g["score"], g["score64"] = edge_weights(g["from"], g["to"], g["sv"], unit, raw,
                                                colnorm if S > 1 else None, S, mode=1)
        g["score_ref"], _ = edge_weights(g["from"], g["to"], g["sv"], unit, raw,
                                         colnorm if S > 1 else None, S, mode=0)
        g["near"] = edge_near(g["from"], g["to"], g["sv"], unit, raw, colnorm if S > 1 else None, S,
                              near_tol)
    else:
        g["near"] = np.zeros(0, np.uint64)
    return g


# ---- primitives used by the pinning tests ---------------------------------

def _rng_lib(which):
    return orc() if which == "orc" else ref()


def wyrand_seq(state, n, which="orc"):
    st = np.array([state], np.uint64)
    out = np.zeros(n, np.uint64)
    getattr(_rng_lib(which), f"{which}_wyrand_seq")(_p(st, u64p), n, _p(out, u64p))
    return out, int(st[0])


def wy_uniform_int_seq(state, bound, n, which="orc"):
    st = np.array([state], np.uint64)
    out = np.zeros(n, np.uint32)
    getattr(_rng_lib(which), f"{which}_wy_uniform_int_seq")(_p(st, u64p), bound, n, _p(out, u32p))
    return out, int(st[0])


def zig_seq(state, n, which="orc"):
    st = np.array([state], np.uint64)
    out = np.zeros(n, np.float64)
    getattr(_rng_lib(which), f"{which}_zig_seq")(_p(st, u64p), n, _p(out, f64p))
    return out, int(st[0])


def ted_seq(c1, c2, c3, state, n, which="orc"):
    st = np.array([state], np.uint64)
    out = np.zeros(n, np.float64)
    getattr(_rng_lib(which), f"{which}_ted_seq")(c1, c2, c3, _p(st, u64p), n, _p(out, f64p))
    return out, int(st[0])


def ted_params(m):
    """TED table rows (boundary, gap, c1, c2, c3) for i < m-1, and 1/firstBoundary."""
    out = np.zeros((m - 1, 5), np.float64)
    fbi = C.c_double()
    orc().orc_ted_params(m, _p(out, f64p), C.byref(fbi))
    return out, fbi.value


def fmix64(x, seed, which="orc"):
    return int(getattr(_rng_lib(which), f"{which}_fmix64")(x, seed))


def pmh_stream(hashes, weights, m, seed=42, which="orc"):
    """Insert (hash, weight) pairs in order into a fresh streaming sketch;
    return (winners u64[m], registers f64[m])."""
    L = _rng_lib(which)
    h = L.orc_pmh_new(m, seed) if which == "orc" else L.ref_pmh_new(m, 4, seed)
    try:
        for x, w in zip(hashes, weights):
            getattr(L, f"{which}_pmh_add_hash")(h, int(x), float(w))
        win = np.zeros(m, np.uint64)
        reg = np.zeros(m, np.float64)
        getattr(L, f"{which}_pmh_get")(h, _p(win, u64p), _p(reg, f64p))
    finally:
        getattr(L, f"{which}_pmh_free")(h)
    return win, reg


def zig_tables_ref():
    x = np.zeros(257, np.float64)
    y = np.zeros(257, np.float64)
    ref().ref_zig_tables(_p(x, f64p), _p(y, f64p))
    return x, y


def kmer4_hist(seq):
    cnt = np.zeros(256, np.uint32)
    base = np.zeros(4, np.uint64)
    tot = C.c_uint64()
    orc().orc_kmer4_hist(seq, len(seq), _p(cnt, u32p), _p(base, u64p), C.byref(tot))
    return cnt, base, tot.value


def kmer4_weights(cnt, base, total, cap=20.0):
    w = np.zeros(256, np.float64)
    c = np.ascontiguousarray(cnt, np.uint32)
    b = np.ascontiguousarray(base, np.uint64)
    orc().orc_kmer4_weights(_p(c, u32p), _p(b, u64p), total, cap, _p(w, f64p))
    return w


def rand_seq(seed, n):
    out = np.zeros(n, np.int32)
    orc().orc_rand_seq(seed, n, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


def mt19937_64_nth(k):
    L = orc()
    L.orc_mt19937_64_nth.restype = C.c_uint64
    L.orc_mt19937_64_nth.argtypes = [C.c_uint64]
    return int(L.orc_mt19937_64_nth(k))


def baseline_draws(n):
    """The 30000 (i, j) draws of estimate_pmh_baseline for n rows."""
    L = orc()
    L.orc_baseline_draws.restype = C.c_uint32
    L.orc_baseline_draws.argtypes = [C.c_uint64, u64p, u64p]
    i = np.zeros(30000, np.uint64)
    j = np.zeros(30000, np.uint64)
    L.orc_baseline_draws(n, _p(i, u64p), _p(j, u64p))
    return i, j


def unit_dot_f(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return np.float32(orc().orc_unit_dot_f(_p(a, f32p), _p(b, f32p), len(a)))


def unit_dot_f_avx512(a, b):
    """The same dot product through AVX-512 intrinsics compiled by the host
    compiler (None when the host lacks AVX-512F)."""
    L = orc()
    L.orc_unit_dot_f_avx512.restype = C.c_int
    L.orc_unit_dot_f_avx512.argtypes = [f32p, f32p, C.c_uint32, f32p]
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = C.c_float()
    if L.orc_unit_dot_f_avx512(_p(a, f32p), _p(b, f32p), len(a), C.byref(out)) != 0:
        return None
    return np.float32(out.value)


def depth_corr(a, b, mode):
    """The edge weight's correlation term: mode 0 in the reference's AVX-512
    order (FP32 chunks, FP64 tail), mode 1 in FP64 throughout."""
    L = orc()
    L.orc_depth_corr.restype = C.c_double
    L.orc_depth_corr.argtypes = [f32p, f32p, C.c_uint32, C.c_int]
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return L.orc_depth_corr(_p(a, f32p), _p(b, f32p), len(a), mode)


def depth_corr_avx512(a, b):
    """depth_corr_fast through AVX-512 intrinsics (None without AVX-512F)."""
    L = orc()
    L.orc_depth_corr_avx512.restype = C.c_int
    L.orc_depth_corr_avx512.argtypes = [f32p, f32p, C.c_uint32, f64p]
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = C.c_double()
    if L.orc_depth_corr_avx512(_p(a, f32p), _p(b, f32p), len(a), C.byref(out)) != 0:
        return None
    return out.value


def sv_from_count(cnt, m, b0):
    return orc().orc_sv_from_count(cnt, m, b0)


def calibrate(rowmax, m, b0, regime, coverage=0.95, rand_seed=1):
    L = orc()
    L.orc_calibrate.restype = C.c_double
    L.orc_calibrate.argtypes = [u32p, C.c_uint64, C.c_uint32, C.c_double, C.c_double, C.c_uint, C.c_int]
    rm = np.ascontiguousarray(rowmax, np.uint32)
    return L.orc_calibrate(_p(rm, u32p), len(rm), m, b0, coverage, rand_seed, regime)


def rows_topk(win16, unit, anynz, S, b0, rows, K=200, thr=None, threads=0):
    """Fused-regime top-K (ids sorted), row maxima and gated-partner counts of
    the given rows against all partners (thr None: the default gate value)."""
    L = orc()
    L.orc_rows_topk.argtypes = [u16p, f32p, u8p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                C.c_uint32, C.c_float, u32p, C.c_uint32, u32p, u32p, u32p, u64p, C.c_int]
    w = np.ascontiguousarray(win16, np.uint16)
    n, m = w.shape
    u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
    nz = np.ascontiguousarray(anynz, np.uint8) if anynz is not None else None
    r = np.ascontiguousarray(rows, np.uint32)
    if thr is None:
        thr = np.float32(0.7 - 1e-4) if (u is not None and S > 2) else -np.inf
    ids = np.zeros((len(r), K), np.uint32)
    ln = np.zeros(len(r), np.uint32)
    rm = np.zeros(len(r), np.uint32)
    gt = np.zeros(len(r), np.uint64)
    L.orc_rows_topk(_p(w, u16p), _p(u, f32p), _p(nz, u8p), n, m, S, b0, K, float(thr), _p(r, u32p), len(r),
                    _p(ids, u32p), _p(ln, u32p), _p(rm, u32p), _p(gt, u64p), threads)
    return ids, ln, rm, gt
