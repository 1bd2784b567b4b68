"""ctypes binding of librbg (include/rbg.h).

Mirrors the reference seams the library replaces:
  build_pmh_winners      (src/rabbitbin.cpp:1436)      -> Context.build_pmh_winners
  estimate_pmh_baseline  (src/rabbitbin.cpp:1340)      -> Context.pmh_baseline
  gen_fused_calib_graph / calibrate_sim_cutoff + build_similarity_graph
                         (src/impl/rb_graph.cpp:501, 1362, 199) -> Context.build_graph
  edge-weight loop       (src/rabbitbin.cpp:5482-5509) -> Context.edge_weights
Errors from the library raise RbgError with the library's message.
"""
import ctypes as C

import numpy as np

from . import LIB_PATH

u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class RbgError(RuntimeError):
    pass


class GraphParams(C.Structure):
    _fields_ = [
        ("m", C.c_uint32),
        ("max_edges", C.c_uint32),
        ("min_edge_weight", C.c_double),
        ("coverage", C.c_double),
        ("min_cv", C.c_double),
        ("neg_depth_thr", C.c_double),
        ("pmh_seed", C.c_uint64),
        ("rand_seed", C.c_uint32),
        ("regime", C.c_int),
        ("flags", C.c_uint32),
        ("near_tol", C.c_double),
        ("semantics", C.c_uint32),
    ]


EDGES_ON_DEVICE = 1
BUILD_CAPTURE = 2

# RBG_SEM_* bits (include/rbg.h): RABBIT_* variables selecting non-default
# reference semantics, which every build refuses.
SEMANTICS = {
    "RABBIT_PMH": 1 << 0, "RABBIT_PMHK": 1 << 1, "RABBIT_PMH_BASE": 1 << 2, "RABBIT_GC_NORM": 1 << 3,
    "RABBIT_GC_NORM_CAP": 1 << 4, "RABBIT_IDF_NORM": 1 << 5, "RABBIT_EXACT_COS": 1 << 6,
    "RABBIT_WIN_BITS": 1 << 7, "RABBIT_LSH": 1 << 8, "RABBIT_PMH_INDEX": 1 << 9,
    "RABBIT_GRAPH_INDEX": 1 << 10, "RABBIT_NO_ABDFIRST": 1 << 11, "RABBIT_MUTUAL_KNN": 1 << 12,
    "RABBIT_W_COMP": 1 << 13, "RABBIT_EDGE_CUT": 1 << 14, "RABBIT_EDGE_POWER": 1 << 15,
    "RABBIT_SNN": 1 << 16, "RABBIT_DEPTH_SIM": 1 << 17, "RABBIT_DEPTH_FUSE": 1 << 18,
    "RABBIT_DEPTH_WJAC_NORM": 1 << 19, "RABBIT_DEPTH_WJAC_CUT": 1 << 20, "RABBIT_DUAL_CONJ": 1 << 21,
    "RABBIT_SNV_EDGE": 1 << 22,
}
RBG_SEM_NO_MUTUAL = SEMANTICS["RABBIT_MUTUAL_KNN"]


class IngestInfo(C.Structure):
    _fields_ = [("records", C.c_uint64), ("n_large", C.c_uint32), ("n_small", C.c_uint32),
                ("large_bases", C.c_uint64), ("names", C.c_void_p), ("name_off", u64p), ("lens", u64p),
                ("small_names", C.c_void_p), ("small_name_off", u64p), ("small_lens", u64p),
                ("S", C.c_uint32), ("depth_rows", C.c_uint64), ("ms_fasta", C.c_double), ("ms_depth", C.c_double)]


class AssemblyIn(C.Structure):
    _fields_ = [("seqs", C.c_char_p), ("offsets", C.POINTER(C.c_uint64)), ("n", C.c_uint32),
                ("depth", C.POINTER(C.c_float)), ("S", C.c_uint32)]


class Schedule(C.Structure):
    _fields_ = [("gate_tiles_per_item", C.c_uint32), ("dense_rows_per_item", C.c_uint32)]


class BatchStats(C.Structure):
    _fields_ = [("candidates", C.c_uint64), ("edges", C.c_uint64), ("ms_total", C.c_double),
                ("replayed", C.c_int)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char_p), ("ms", C.c_double)]


class GraphResult(C.Structure):
    _fields_ = [
        ("n_edges", C.c_uint64),
        ("from_", u32p),
        ("to", u32p),
        ("scomp", f32p),
        ("score", f32p),
        ("sim_cutoff", C.c_double),
        ("b0", C.c_double),
        ("regime", C.c_int),
        ("gated_pairs", C.c_uint64),
        ("ms_sketch", C.c_double),
        ("ms_pairs", C.c_double),
        ("ms_select", C.c_double),
        ("ms_edges", C.c_double),
        ("ms_total", C.c_double),
        ("ms_k_pair", C.c_double),
        ("ms_k_compare", C.c_double),
        ("ms_k_select", C.c_double),
        ("n_near", C.c_uint64),
        ("near_edge", u64p),
        ("near_score64", f64p),
        ("n_kernels", C.c_uint32),
        ("kernels", KernelTime * 16),
        ("candidates", C.c_uint64),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_step_requests", C.c_uint32), ("max_step_rows", C.c_uint32),
                ("params", GraphParams), ("schedule", Schedule)]


class RequestInfo(C.Structure):
    _fields_ = [("session", C.c_uint64), ("n", C.c_uint32), ("n_new", C.c_uint32), ("incremental", C.c_int),
                ("status", C.c_int), ("step", C.c_uint64), ("step_requests", C.c_uint32),
                ("pass_pairs", C.c_uint64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("latency_ms", C.c_double), ("error", C.c_char * 160)]


class ServerStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("requests", C.c_uint64), ("sessions", C.c_uint64),
                ("pending", C.c_uint64), ("running", C.c_uint64), ("graph_pairs", C.c_uint64),
                ("pass_pairs", C.c_uint64), ("resident_bytes", C.c_uint64), ("workspace_bytes", C.c_uint64),
                ("budget", C.c_uint64), ("busy_ms", C.c_double), ("upload_ms", C.c_double),
                ("build_ms", C.c_double), ("result_ms", C.c_double), ("save_ms", C.c_double)]


class GraphDiag(C.Structure):
    _fields_ = [
        ("topk_id", u32p),
        ("topk_cnt", u16p),
        ("topk_len", u32p),
        ("rowmax_cnt", u32p),
        ("win16", u16p),
    ]


class CacheFile(C.Structure):
    _fields_ = [("version", C.c_uint32), ("nobs", C.c_uint64), ("nobs1", C.c_uint64), ("samples", C.c_uint64),
                ("seed", C.c_uint64), ("total_size", C.c_uint64), ("total_size1", C.c_uint64),
                ("sim_cutoff", C.c_double), ("b0", C.c_double), ("min_contig", C.c_uint64),
                ("min_small_contig", C.c_uint64), ("has_depth", C.c_int),
                ("names", C.c_void_p), ("name_off", u64p), ("small_names", C.c_void_p), ("small_name_off", u64p),
                ("seq_lens", u64p), ("small_seq_lens", u64p),
                ("depth_rows", C.c_uint64), ("depth_cols", C.c_uint64), ("depth", f32p),
                ("small_depth_rows", C.c_uint64), ("small_depth_cols", C.c_uint64), ("small_depth", f32p),
                ("n_large_means", C.c_uint64), ("large_means", f32p),
                ("n_depth_raw", C.c_uint64), ("depth_raw", f32p),
                ("n_colnorm", C.c_uint64), ("colnorm", f64p),
                ("n_edges", C.c_uint64), ("from_", u32p), ("to", u32p), ("scomp", f32p),
                ("has_snv", C.c_int), ("n_snv", C.c_uint64), ("snv_pi", f32p), ("snv_nsnv", u32p),
                ("snv_cov", u32p)]


_lib = None


def lib():
    """Load librbg.so; raise if it is not built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise RbgError(f"librbg.so is not loadable ({e}); build it with __graft_entry__.build()")
    vp = C.c_void_p
    L.rbg_open.argtypes = [C.c_int, C.POINTER(vp)]
    L.rbg_close.argtypes = [vp]
    L.rbg_last_error.argtypes = [vp]
    L.rbg_last_error.restype = C.c_char_p
    L.rbg_default_params.argtypes = [C.POINTER(GraphParams)]
    L.rbg_sketch_winners.argtypes = [vp, C.c_char_p, u64p, C.c_uint32, C.c_uint32, C.c_uint64, u32p, u16p]
    L.rbg_depth_features.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_double, f32p, f32p, f32p, f64p, u8p]
    L.rbg_build_graph.argtypes = [vp, C.c_char_p, u64p, C.c_uint32, f32p, C.c_uint32,
                                  C.POINTER(GraphParams), C.POINTER(GraphResult), C.POINTER(GraphDiag)]
    L.rbg_build_graph_from_sketches.argtypes = [vp, u16p, C.c_uint32, C.c_uint32, f64p, f32p, f32p, f64p, u8p,
                                                C.c_uint32, C.POINTER(GraphParams), C.POINTER(GraphResult),
                                                C.POINTER(GraphDiag)]
    L.rbg_asm_create.argtypes = [vp, C.c_char_p, u64p, C.c_uint32, f32p, C.c_uint32, C.c_double, C.POINTER(vp)]
    L.rbg_asm_free.argtypes = [vp]
    L.rbg_asm_copy_edges.argtypes = [vp, vp, u32p, u32p, f32p, f32p]
    L.rbg_sketch_tables.argtypes = [C.c_uint32, C.c_uint64, f64p, u64p, u8p]
    L.rbg_calibrate.argtypes = [u32p, C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_uint32, C.c_int, f64p]
    L.rbg_asm_build_graph.argtypes = [vp, vp, C.POINTER(GraphParams), C.POINTER(GraphResult), C.POINTER(GraphDiag)]
    L.rbg_graph_result_free.argtypes = [C.POINTER(GraphResult)]
    L.rbg_edge_weights.argtypes = [vp, u32p, u32p, f32p, C.c_uint64, f32p, f32p, f64p, C.c_uint32,
                                   C.c_uint32, C.c_double, C.c_double, C.c_double, f32p, f64p]
    L.rbg_pmh_baseline.argtypes = [vp, u16p, C.c_uint32, C.c_uint32, f64p]
    L.rbg_cache_write.argtypes = [C.c_char_p, C.POINTER(CacheFile)]
    L.rbg_cache_read.argtypes = [C.c_char_p, C.POINTER(CacheFile)]
    L.rbg_cache_free.argtypes = [C.POINTER(CacheFile)]
    L.rbg_label_propagation.argtypes = [u32p, u32p, f32p, C.c_uint64, C.c_uint32, C.c_uint64, u32p,
                                        C.POINTER(C.c_uint32)]
    L.rbg_lpa_node_order.argtypes = [C.c_uint32, C.c_uint64, u32p]
    i32p = C.POINTER(C.c_int32)
    L.rbg_recruit_frozen_cores.argtypes = [vp, C.c_uint32, C.c_uint32, f32p, u64p, C.c_uint32, f32p, u64p,
                                           C.c_uint32, i32p, u64p, u64p, C.c_uint64, C.c_uint64, i32p, i32p,
                                           C.POINTER(RecruitStats)]
    L.rbg_env_semantics.argtypes = [C.c_char_p, C.c_size_t]
    L.rbg_env_semantics.restype = C.c_uint32
    L.rbg_batch_create.argtypes = [vp, C.POINTER(AssemblyIn), C.c_uint32, C.c_double, C.POINTER(vp)]
    L.rbg_batch_set_schedule.argtypes = [vp, C.POINTER(Schedule)]
    L.rbg_batch_build.argtypes = [vp, vp, C.POINTER(GraphParams), C.POINTER(BatchStats)]
    L.rbg_batch_result.argtypes = [vp, vp, C.c_uint32, C.POINTER(GraphResult)]
    L.rbg_batch_autotune.argtypes = [vp, vp, C.POINTER(GraphParams), C.c_uint32, C.POINTER(Schedule),
                                     C.POINTER(C.c_double)]
    L.rbg_ingest_fasta.argtypes = [vp, C.c_char_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(vp)]
    L.rbg_ingest_depth.argtypes = [vp, vp, C.c_char_p, C.c_uint64, C.c_uint32, C.c_int, C.c_uint64]
    L.rbg_ingest_get.argtypes = [vp, C.POINTER(IngestInfo)]
    L.rbg_ingest_copy.argtypes = [vp, vp, C.c_char_p, u64p, f32p, f32p]
    L.rbg_ingest_assembly.argtypes = [vp, vp, C.c_double, C.POINTER(vp)]
    L.rbg_ingest_free.argtypes = [vp]
    L.rbg_server_default_config.argtypes = [C.POINTER(ServerConfig)]
    L.rbg_server_start.argtypes = [C.c_int, C.POINTER(ServerConfig), C.POINTER(vp)]
    L.rbg_server_stop.argtypes = [vp]
    L.rbg_server_last_error.argtypes = [vp]
    L.rbg_server_last_error.restype = C.c_char_p
    L.rbg_server_submit.argtypes = [vp, C.POINTER(AssemblyIn), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.rbg_server_extend.argtypes = [vp, C.c_uint64, C.POINTER(AssemblyIn), C.POINTER(C.c_uint64)]
    L.rbg_server_wait.argtypes = [vp, C.c_uint64, C.c_double, C.POINTER(GraphResult), C.POINTER(RequestInfo)]
    L.rbg_server_release.argtypes = [vp, C.c_uint64]
    L.rbg_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


_CACHE_SCALARS = ("nobs", "nobs1", "samples", "seed", "total_size", "total_size1", "sim_cutoff", "b0",
                  "min_contig", "min_small_contig", "has_depth")


def _names_buffer(names):
    enc = [x.encode() if isinstance(x, str) else bytes(x) for x in names]
    off = np.zeros(len(enc) + 1, np.uint64)
    off[1:] = np.cumsum([len(x) for x in enc])
    return np.frombuffer(b"".join(enc) or b"\0", np.uint8).copy(), off


def write_cache(path, **kw):
    """rbg_cache_write: an RBCACHE1 v3 file from keyword fields (the scalars
    of rbg_cache, names/small_names as lists of str, seq_lens/small_seq_lens,
    depth/small_depth as 2-D float32, large_means, depth_raw, colnorm, and the
    graph as from_/to/scomp)."""
    c = CacheFile()
    keep = []
    for k in _CACHE_SCALARS:
        if k in kw:
            setattr(c, k, kw[k])

    def arr(name, dt, ptr_t):
        a = kw.get(name)
        if a is None:
            return None, 0
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data_as(ptr_t), a.size

    for key, off_key, n_key in (("names", "name_off", "nobs"), ("small_names", "small_name_off", "nobs1")):
        if kw.get(key) is not None:
            buf, off = _names_buffer(kw[key])
            keep += [buf, off]
            setattr(c, key, buf.ctypes.data)
            setattr(c, off_key, off.ctypes.data_as(u64p))
    c.seq_lens, _ = arr("seq_lens", np.uint64, u64p)
    c.small_seq_lens, _ = arr("small_seq_lens", np.uint64, u64p)
    for key, rk, ck in (("depth", "depth_rows", "depth_cols"), ("small_depth", "small_depth_rows", "small_depth_cols")):
        if kw.get(key) is not None:
            a = np.ascontiguousarray(kw[key], np.float32)
            keep.append(a)
            setattr(c, key, a.ctypes.data_as(f32p))
            setattr(c, rk, a.shape[0])
            setattr(c, ck, a.shape[1])
    c.large_means, c.n_large_means = arr("large_means", np.float32, f32p)
    c.depth_raw, c.n_depth_raw = arr("depth_raw", np.float32, f32p)
    c.colnorm, c.n_colnorm = arr("colnorm", np.float64, f64p)
    c.from_, c.n_edges = arr("from_", np.uint32, u32p)
    c.to, _ = arr("to", np.uint32, u32p)
    c.scomp, _ = arr("scomp", np.float32, f32p)
    st = lib().rbg_cache_write(path.encode(), C.byref(c))
    if st != 0:
        raise RbgError(f"rbg_cache_write({path}): status {st}")


def read_cache(path):
    """rbg_cache_read: the fields of an RBCACHE1 file (versions 1-3) as a dict."""
    c = CacheFile()
    st = lib().rbg_cache_read(path.encode(), C.byref(c))
    if st != 0:
        raise RbgError(f"rbg_cache_read({path}): status {st}")
    try:
        out = {k: getattr(c, k) for k in ("version",) + _CACHE_SCALARS}

        def arr(ptr, n):
            return np.ctypeslib.as_array(ptr, (n,)).copy() if (n and ptr) else None

        def names(buf, off, n):
            if not buf or not n:
                return None
            o = np.ctypeslib.as_array(off, (n + 1,)).copy()
            raw = C.string_at(buf, int(o[-1]))
            return [raw[o[i]:o[i + 1]].decode() for i in range(n)]

        out["names"] = names(c.names, c.name_off, c.nobs)
        out["small_names"] = names(c.small_names, c.small_name_off, c.nobs1)
        out["seq_lens"] = arr(c.seq_lens, c.nobs)
        out["small_seq_lens"] = arr(c.small_seq_lens, c.nobs1)
        d = arr(c.depth, c.depth_rows * c.depth_cols)
        out["depth"] = None if d is None else d.reshape(c.depth_rows, c.depth_cols)
        d = arr(c.small_depth, c.small_depth_rows * c.small_depth_cols)
        out["small_depth"] = None if d is None else d.reshape(c.small_depth_rows, c.small_depth_cols)
        out["large_means"] = arr(c.large_means, c.n_large_means)
        out["depth_raw"] = arr(c.depth_raw, c.n_depth_raw)
        out["colnorm"] = arr(c.colnorm, c.n_colnorm)
        out["from_"] = arr(c.from_, c.n_edges)
        out["to"] = arr(c.to, c.n_edges)
        out["scomp"] = arr(c.scomp, c.n_edges)
        out["n_edges"] = c.n_edges
        return out
    finally:
        lib().rbg_cache_free(C.byref(c))


def label_propagation(frm, to, score, n, seed=1):
    """rbg_label_propagation: (membership uint32[n], rounds)."""
    frm = np.ascontiguousarray(frm, np.uint32)
    to = np.ascontiguousarray(to, np.uint32)
    sc = np.ascontiguousarray(score, np.float32)
    mem = np.zeros(n, np.uint32)
    rounds = C.c_uint32()
    st = lib().rbg_label_propagation(_p(frm, u32p), _p(to, u32p), _p(sc, f32p), len(frm), n, seed, _p(mem, u32p),
                                     C.byref(rounds))
    if st != 0:
        raise RbgError(f"rbg_label_propagation: status {st}")
    return mem, rounds.value


def lpa_node_order(n, seed=1):
    out = np.zeros(n, np.uint32)
    st = lib().rbg_lpa_node_order(n, seed, _p(out, u32p))
    if st != 0:
        raise RbgError(f"rbg_lpa_node_order: status {st}")
    return out


def env_semantics():
    """rbg_env_semantics: (mask, comma list of the RABBIT_* variables that
    select non-default reference semantics in this process's environment)."""
    buf = C.create_string_buffer(1024)
    mask = lib().rbg_env_semantics(buf, len(buf))
    return mask, buf.value.decode()


def default_params(**kw):
    p = GraphParams()
    lib().rbg_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class _Seqs:
    """Contig bytes passed by address (no copy); keeps the array alive."""

    def __init__(self, seqs):
        if isinstance(seqs, (bytes, bytearray)):
            seqs = np.frombuffer(bytes(seqs), dtype=np.uint8)
        self.arr = np.ascontiguousarray(seqs, dtype=np.uint8)
        if self.arr.size == 0:
            self.arr = np.zeros(1, np.uint8)
        self.ptr = C.c_char_p(self.arr.ctypes.data)

    @property
    def _as_parameter_(self):
        return self.ptr


class _ResultOwner:
    """A graph result's host arrays, freed (rbg_graph_result_free) once no
    array built on them remains: the edge arrays are handed out without a
    copy."""

    def __init__(self, L, res):
        self._L, self._res = L, res

    def __del__(self):
        self._L.rbg_graph_result_free(C.byref(self._res))


def _owned_array(owner, ptr, count):
    """numpy view of count elements at the ctypes pointer ptr that keeps owner
    alive."""
    buf = (ptr._type_ * count).from_address(C.cast(ptr, C.c_void_p).value)
    buf._owner = owner
    return np.frombuffer(buf, dtype=np.dtype(ptr._type_))


def _as_seqs(seqs, offsets):
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    return _Seqs(seqs), off


class RecruitStats(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in ("cores", "long_assigned", "long_ambiguous", "long_unmatched",
                                         "short_assigned", "short_ambiguous", "short_unmatched")]


def bins_csr(bins):
    """(ids int32, offsets uint64, members uint64) of a {bin id: [members]}
    mapping, bins in the mapping's order."""
    ids = np.array(list(bins.keys()), dtype=np.int32)
    sizes = [len(v) for v in bins.values()]
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes, dtype=np.uint64)
    mem = np.concatenate([np.asarray(v, np.uint64) for v in bins.values()]) if sizes else np.zeros(0, np.uint64)
    return ids, off, np.ascontiguousarray(mem, np.uint64)


def apply_recruitment(bins, assign_large, assign_small, n_large):
    """The bins after recruitment, as the reference appends: the assigned
    large contigs in ascending order, then the small ones (n_large + s)."""
    out = {k: list(v) for k, v in bins.items()}
    for c in np.nonzero(np.asarray(assign_large) >= 0)[0]:
        out[int(assign_large[c])].append(int(c))
    for s in np.nonzero(np.asarray(assign_small) >= 0)[0]:
        out[int(assign_small[s])].append(int(n_large + s))
    return out


class Context:
    """One HIP device + stream (rbg_open)."""

    def __init__(self, device=0):
        self._L = lib()
        self._h = C.c_void_p()
        st = self._L.rbg_open(device, C.byref(self._h))
        if st != 0:
            raise RbgError(f"rbg_open({device}) failed with status {st}")

    def close(self):
        if self._h:
            self._L.rbg_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st != 0:
            raise RbgError(f"{what}: status {st}: {self._L.rbg_last_error(self._h).decode()}")

    def build_pmh_winners(self, seqs, offsets, m=500, seed=42):
        buf, off = _as_seqs(seqs, offsets)
        n = len(off) - 1
        w32 = np.zeros((n, m), np.uint32)
        w16 = np.zeros((n, m), np.uint16)
        self._check(self._L.rbg_sketch_winners(self._h, buf, _p(off, u64p), n, m, seed,
                                               _p(w32, u32p), _p(w16, u16p)), "rbg_sketch_winners")
        return w32, w16

    def recruit_frozen_cores(self, S, unit_large, len_large, depth_small, len_small, bins, core_floor,
                             min_cs=10):
        """rbg_recruit_frozen_cores: (assign_large int32, assign_small int32,
        stats dict).  bins: {bin id: [members]} (large c, small n_large + s)."""
        ul = np.ascontiguousarray(unit_large, np.float32).reshape(-1)
        ll = np.ascontiguousarray(len_large, np.uint64)
        ds = np.ascontiguousarray(depth_small, np.float32).reshape(-1)
        ls = np.ascontiguousarray(len_small, np.uint64)
        nl, ns = len(ll), len(ls)
        ids, off, mem = bins_csr(bins)
        al = np.zeros(nl, np.int32)
        as_ = np.zeros(ns, np.int32)
        stt = RecruitStats()
        i32p = C.POINTER(C.c_int32)
        self._check(self._L.rbg_recruit_frozen_cores(self._h, S, nl, _p(ul, f32p), _p(ll, u64p), ns, _p(ds, f32p),
                                                     _p(ls, u64p), len(ids), _p(ids, i32p), _p(off, u64p),
                                                     _p(mem, u64p), core_floor, min_cs, _p(al, i32p),
                                                     _p(as_, i32p), C.byref(stt)), "rbg_recruit_frozen_cores")
        return al, as_, {f: getattr(stt, f) for f, _ in RecruitStats._fields_}

    def pmh_baseline(self, win16):
        w = np.ascontiguousarray(win16, dtype=np.uint16)
        b0 = C.c_double()
        self._check(self._L.rbg_pmh_baseline(self._h, _p(w, u16p), w.shape[0], w.shape[1], C.byref(b0)),
                    "rbg_pmh_baseline")
        return b0.value

    def _result(self, res, n, K, diag_arrays):
        ne = res.n_edges
        host = bool(res.from_)
        owner = _ResultOwner(self._L, res)

        def arr(ptr, dt):
            return _owned_array(owner, ptr, ne) if (ne and ptr) else np.zeros(0, dt)

        out = {
            "n_edges": ne,
            "from": arr(res.from_, np.uint32) if host else None,
            "to": arr(res.to, np.uint32) if host else None,
            "scomp": arr(res.scomp, np.float32) if host else None,
            "score": _owned_array(owner, res.score, ne) if (ne and res.score) else None,
            "sim_cutoff": res.sim_cutoff, "b0": res.b0, "regime": res.regime,
            "gated_pairs": res.gated_pairs,
            "candidates": res.candidates,
            "near_edge": (np.ctypeslib.as_array(res.near_edge, (res.n_near,)).copy() if res.n_near
                          else np.zeros(0, np.uint64)),
            "near_score64": (np.ctypeslib.as_array(res.near_score64, (res.n_near,)).copy() if res.n_near
                             else np.zeros(0, np.float64)),
            "ms": {"sketch": res.ms_sketch, "pairs": res.ms_pairs, "select": res.ms_select,
                   "edges": res.ms_edges, "total": res.ms_total, "k_pair": res.ms_k_pair,
                   "k_compare": res.ms_k_compare, "k_select": res.ms_k_select},
            "kernels": [(res.kernels[k].name.decode(), res.kernels[k].ms) for k in range(res.n_kernels)],
        }
        if diag_arrays is not None:
            out.update(diag_arrays)
        return out  # the library's arrays are freed with the last array built on them

    def _diag(self, n, K, m, want):
        if not want:
            return None, None
        arrs = {"topk_id": np.zeros((n, K), np.uint32), "topk_cnt": np.zeros((n, K), np.uint16),
                "topk_len": np.zeros(n, np.uint32), "rowmax_cnt": np.zeros(n, np.uint32),
                "win16": np.zeros((n, m), np.uint16)}
        d = GraphDiag(_p(arrs["topk_id"], u32p), _p(arrs["topk_cnt"], u16p), _p(arrs["topk_len"], u32p),
                      _p(arrs["rowmax_cnt"], u32p), _p(arrs["win16"], u16p))
        return d, arrs

    def build_graph(self, seqs, offsets, depth=None, params=None, diag=False):
        buf, off = _as_seqs(seqs, offsets)
        n = len(off) - 1
        p = params or default_params()
        if depth is None:
            dep, S = None, 0
        else:
            dep = np.ascontiguousarray(depth, dtype=np.float32)
            S = dep.shape[1]
        d, arrs = self._diag(n, p.max_edges, p.m, diag)
        res = GraphResult()
        self._check(self._L.rbg_build_graph(self._h, buf, _p(off, u64p), n, _p(dep, f32p), S, C.byref(p),
                                            C.byref(res), C.byref(d) if d else None), "rbg_build_graph")
        return self._result(res, n, p.max_edges, arrs)

    def build_graph_from_sketches(self, win16, unit=None, raw=None, colnorm=None, anynz=None, b0=None,
                                  params=None, diag=False):
        """rbg_build_graph_from_sketches: the graph from the caller's g_win16,
        depth rows, g_anynz and (optionally) b0."""
        w = np.ascontiguousarray(win16, np.uint16)
        n, m = w.shape
        p = params or default_params()
        u = None if unit is None else np.ascontiguousarray(unit, np.float32)
        r = None if raw is None else np.ascontiguousarray(raw, np.float32)
        cn = None if colnorm is None else np.ascontiguousarray(colnorm, np.float64)
        nz = None if anynz is None else np.ascontiguousarray(anynz, np.uint8)
        S = 0 if u is None else u.shape[1]
        b0v = None if b0 is None else C.c_double(b0)
        d, arrs = self._diag(n, p.max_edges, p.m, diag)
        if d is not None:
            d.win16 = None
            arrs.pop("win16")
        res = GraphResult()
        self._check(self._L.rbg_build_graph_from_sketches(self._h, _p(w, u16p), n, m,
                                                          C.byref(b0v) if b0v is not None else None,
                                                          _p(u, f32p), _p(r, f32p), _p(cn, f64p), _p(nz, u8p), S,
                                                          C.byref(p), C.byref(res), C.byref(d) if d else None),
                    "rbg_build_graph_from_sketches")
        return self._result(res, n, p.max_edges, arrs)

    def edge_weights(self, frm, to, scomp, unit, raw, colnorm, S, w_comp, min_edge_weight=0.7,
                     neg_depth_thr=-0.3):
        frm = np.ascontiguousarray(frm, np.uint32)
        to = np.ascontiguousarray(to, np.uint32)
        sc = np.ascontiguousarray(scomp, np.float32)
        u = np.ascontiguousarray(unit, np.float32) if unit is not None else None
        r = np.ascontiguousarray(raw, np.float32) if raw is not None else None
        cn = np.ascontiguousarray(colnorm, np.float64) if colnorm is not None else None
        n = u.shape[0] if u is not None else int(max(frm.max(initial=0), to.max(initial=0)) + 1)
        w = np.zeros(len(frm), np.float32)
        w64 = np.zeros(len(frm), np.float64)
        self._check(self._L.rbg_edge_weights(self._h, _p(frm, u32p), _p(to, u32p), _p(sc, f32p), len(frm),
                                             _p(u, f32p), _p(r, f32p), _p(cn, f64p), n, S, w_comp,
                                             min_edge_weight, neg_depth_thr, _p(w, f32p), _p(w64, f64p)),
                    "rbg_edge_weights")
        return w, w64


class Assembly:
    """Device-resident assembly (rbg_asm_create): contigs + depth features in HBM."""

    def __init__(self, ctx, seqs, offsets, depth=None, min_cv=1.0):
        self.ctx = ctx
        buf, off = _as_seqs(seqs, offsets)
        self.n = len(off) - 1
        dep = None if depth is None else np.ascontiguousarray(depth, np.float32)
        self.S = 0 if dep is None else dep.shape[1]
        self._h = C.c_void_p()
        ctx._check(ctx._L.rbg_asm_create(ctx._h, buf, _p(off, u64p), self.n, _p(dep, f32p), self.S, min_cv,
                                         C.byref(self._h)), "rbg_asm_create")

    def build_graph(self, params=None, diag=False, on_device=False, capture=False):
        """on_device=True leaves the edge list in HBM (fetch with copy_edges);
        capture=True replays a captured graph of the build after the first call."""
        p = GraphParams.from_buffer_copy(params or default_params())
        if on_device:
            p.flags |= EDGES_ON_DEVICE
        if capture:
            p.flags |= BUILD_CAPTURE
        d, arrs = self.ctx._diag(self.n, p.max_edges, p.m, diag)
        res = GraphResult()
        self.ctx._check(self.ctx._L.rbg_asm_build_graph(self.ctx._h, self._h, C.byref(p), C.byref(res),
                                                        C.byref(d) if d else None), "rbg_asm_build_graph")
        return self.ctx._result(res, self.n, p.max_edges, arrs)

    def copy_edges(self, ne):
        frm = np.zeros(ne, np.uint32)
        to = np.zeros(ne, np.uint32)
        sc = np.zeros(ne, np.float32)
        score = np.zeros(ne, np.float32) if self.S else None
        self.ctx._check(self.ctx._L.rbg_asm_copy_edges(self.ctx._h, self._h, _p(frm, u32p), _p(to, u32p),
                                                       _p(sc, f32p), _p(score, f32p)), "rbg_asm_copy_edges")
        return frm, to, sc, score

    def free(self):
        if self._h:
            self.ctx._L.rbg_asm_free(self._h)
            self._h = C.c_void_p()


class Ingest:
    """FASTA and depth file parsed on the device (rbg_ingest_fasta /
    rbg_ingest_depth); the large contigs stay in HBM."""

    def __init__(self, ctx, fasta, min_contig=2500, min_small_contig=1000):
        self.ctx = ctx
        data = bytes(fasta)
        self.min_small_contig = min_small_contig
        self._h = C.c_void_p()
        ctx._check(ctx._L.rbg_ingest_fasta(ctx._h, data, len(data), min_contig, min_small_contig,
                                           C.byref(self._h)), "rbg_ingest_fasta")

    def depth(self, data, S=0, cv_ext=False, min_depth_len=None):
        """Parse the depth file; min_depth_len defaults to min_small_contig
        (the reference's prefilter)."""
        data = bytes(data)
        mdl = self.min_small_contig if min_depth_len is None else min_depth_len
        self.ctx._check(self.ctx._L.rbg_ingest_depth(self.ctx._h, self._h, data, len(data), S, int(cv_ext), mdl),
                        "rbg_ingest_depth")

    def info(self):
        i = IngestInfo()
        st = self.ctx._L.rbg_ingest_get(self._h, C.byref(i))
        if st != 0:
            raise RbgError(f"rbg_ingest_get: status {st}")

        def names(ptr, off, n):
            if n == 0:
                return []
            o = np.ctypeslib.as_array(off, (n + 1,)).copy()
            blob = C.string_at(ptr, int(o[-1])) if o[-1] else b""
            return [blob[o[k]:o[k + 1]] for k in range(n)]

        def arr(p, n):
            return np.ctypeslib.as_array(p, (n,)).copy() if n else np.zeros(0, np.uint64)

        return {"records": i.records, "n_large": i.n_large, "n_small": i.n_small, "large_bases": i.large_bases,
                "names": names(i.names, i.name_off, i.n_large), "lens": arr(i.lens, i.n_large),
                "small_names": names(i.small_names, i.small_name_off, i.n_small),
                "small_lens": arr(i.small_lens, i.n_small), "S": i.S, "depth_rows": i.depth_rows,
                "ms_fasta": i.ms_fasta, "ms_depth": i.ms_depth}

    def copy(self):
        """(seqs uint8, offsets uint64[n+1], depth float32[n, S] or None, small_depth or None)."""
        inf = self.info()
        n, ns, S = inf["n_large"], inf["n_small"], inf["S"]
        seqs = np.zeros(max(int(inf["large_bases"]), 1), np.uint8)
        off = np.zeros(n + 1, np.uint64)
        dep = np.zeros((n, S), np.float32) if S else None
        sdep = np.zeros((ns, S), np.float32) if S else None
        self.ctx._check(self.ctx._L.rbg_ingest_copy(self.ctx._h, self._h, C.c_char_p(seqs.ctypes.data),
                                                    _p(off, u64p), _p(dep, f32p), _p(sdep, f32p)),
                        "rbg_ingest_copy")
        return seqs[:int(inf["large_bases"])], off, dep, sdep

    def assembly(self, min_cv=1.0):
        """A resident Assembly of the large contigs and their depth rows."""
        a = Assembly.__new__(Assembly)
        a.ctx = self.ctx
        inf = self.info()
        a.n, a.S = inf["n_large"], inf["S"]
        a._h = C.c_void_p()
        self.ctx._check(self.ctx._L.rbg_ingest_assembly(self.ctx._h, self._h, min_cv, C.byref(a._h)),
                        "rbg_ingest_assembly")
        return a

    def free(self):
        if self._h:
            self.ctx._L.rbg_ingest_free(self._h)
            self._h = C.c_void_p()


class Batch:
    """Many assemblies resident together (rbg_batch_create); one build runs every
    stage as one launch over all of them.  Each element of `assemblies` is
    (seqs, offsets, depth-or-None)."""

    def __init__(self, ctx, assemblies, min_cv=1.0):
        self.ctx = ctx
        self._keep = []
        ins = (AssemblyIn * len(assemblies))()
        self.sizes = []
        for k, (seqs, offsets, depth) in enumerate(assemblies):
            buf, off = _as_seqs(seqs, offsets)
            dep = None if depth is None else np.ascontiguousarray(depth, np.float32)
            S = 0 if dep is None else dep.shape[1]
            self._keep += [buf, off, dep]
            ins[k] = AssemblyIn(buf.ptr, _p(off, u64p), len(off) - 1, _p(dep, f32p), S)
            self.sizes.append(len(off) - 1)
        self._h = C.c_void_p()
        ctx._check(ctx._L.rbg_batch_create(ctx._h, ins, len(assemblies), min_cv, C.byref(self._h)),
                   "rbg_batch_create")
        self._keep = None  # the library copied the inputs

    def set_schedule(self, gate_tiles_per_item=16, dense_rows_per_item=64):
        st = self.ctx._L.rbg_batch_set_schedule(self._h, C.byref(Schedule(gate_tiles_per_item,
                                                                          dense_rows_per_item)))
        if st != 0:
            raise RbgError(f"rbg_batch_set_schedule: status {st}")

    def build(self, params=None, capture=False):
        p = GraphParams.from_buffer_copy(params or default_params())
        if capture:
            p.flags |= BUILD_CAPTURE
        stats = BatchStats()
        self.ctx._check(self.ctx._L.rbg_batch_build(self.ctx._h, self._h, C.byref(p), C.byref(stats)),
                        "rbg_batch_build")
        return {"candidates": stats.candidates, "edges": stats.edges, "ms": stats.ms_total,
                "replayed": bool(stats.replayed)}

    def autotune(self, params=None, reps=2):
        """rbg_batch_autotune: ((gate_tiles_per_item, dense_rows_per_item), ms) of the
        fastest schedule; the batch keeps it and is left built under it."""
        p = GraphParams.from_buffer_copy(params or default_params())
        best, ms = Schedule(), C.c_double()
        self.ctx._check(self.ctx._L.rbg_batch_autotune(self.ctx._h, self._h, C.byref(p), reps, C.byref(best),
                                                       C.byref(ms)), "rbg_batch_autotune")
        return (best.gate_tiles_per_item, best.dense_rows_per_item), ms.value

    def result(self, index):
        res = GraphResult()
        self.ctx._check(self.ctx._L.rbg_batch_result(self.ctx._h, self._h, index, C.byref(res)),
                        "rbg_batch_result")
        return self.ctx._result(res, self.sizes[index], 0, None)

    def free(self):
        if self._h:
            self.ctx._L.rbg_asm_free(self._h)
            self._h = C.c_void_p()


TIMEOUT = -6


class Server:
    """The served form (rbg_server_*): sessions of device-resident assemblies,
    served by a worker thread in batched steps under an HBM budget.

    submit(seqs, offsets, depth) starts a session; extend(session, ...) adds
    contigs to it; wait(request) returns (result dict, info dict).  Input
    arrays are kept referenced here until their request has been waited for
    (the library reads them while the request runs)."""

    def __init__(self, device=0, hbm_budget=0, max_step_requests=64, max_step_rows=1 << 22, params=None,
                 schedule=None):
        self._L = lib()
        cfg = ServerConfig()
        self._L.rbg_server_default_config(C.byref(cfg))
        cfg.hbm_budget = hbm_budget
        cfg.max_step_requests = max_step_requests
        cfg.max_step_rows = max_step_rows
        if params is not None:
            cfg.params = params
        if schedule is not None:
            cfg.schedule = Schedule(*schedule)
        self.params = GraphParams.from_buffer_copy(cfg.params)
        self._h = C.c_void_p()
        st = self._L.rbg_server_start(device, C.byref(cfg), C.byref(self._h))
        if st != 0:
            raise RbgError(f"rbg_server_start: status {st}")
        self._inputs = {}

    def _err(self):
        return self._L.rbg_server_last_error(self._h).decode()

    @staticmethod
    def _in(seqs, offsets, depth):
        buf, off = _as_seqs(seqs, offsets)
        dep = None if depth is None else np.ascontiguousarray(depth, np.float32)
        S = 0 if dep is None else dep.shape[1]
        return AssemblyIn(buf.ptr, _p(off, u64p), len(off) - 1, _p(dep, f32p), S), (buf, off, dep)

    def submit(self, seqs, offsets, depth=None):
        """New session; returns (session, request)."""
        a, keep = self._in(seqs, offsets, depth)
        ses, req = C.c_uint64(), C.c_uint64()
        st = self._L.rbg_server_submit(self._h, C.byref(a), C.byref(ses), C.byref(req))
        if st != 0:
            raise RbgError(f"rbg_server_submit: status {st}: {self._err()}")
        self._inputs[req.value] = keep
        return ses.value, req.value

    def extend(self, session, seqs, offsets, depth=None):
        """Contigs appended to the session's assembly; returns the request."""
        a, keep = self._in(seqs, offsets, depth)
        req = C.c_uint64()
        st = self._L.rbg_server_extend(self._h, session, C.byref(a), C.byref(req))
        if st != 0:
            raise RbgError(f"rbg_server_extend: status {st}: {self._err()}")
        self._inputs[req.value] = keep
        return req.value

    def wait(self, request, timeout=-1.0):
        res, info = GraphResult(), RequestInfo()
        st = self._L.rbg_server_wait(self._h, request, timeout, C.byref(res), C.byref(info))
        if st == TIMEOUT:
            raise TimeoutError(f"request {request} not done within {timeout} s")
        self._inputs.pop(request, None)
        d = {k: getattr(info, k) for k, _ in RequestInfo._fields_ if k != "error"}
        d["error"] = info.error.decode()
        if st != 0:
            raise RbgError(f"request {request}: status {st}: {d['error']}")
        ne = res.n_edges

        def arr(ptr):
            return np.ctypeslib.as_array(ptr, (ne,)).copy() if (ne and ptr) else None

        out = {"n_edges": ne, "from": arr(res.from_), "to": arr(res.to), "scomp": arr(res.scomp),
               "score": arr(res.score), "sim_cutoff": res.sim_cutoff, "b0": res.b0, "regime": res.regime,
               "gated_pairs": res.gated_pairs,
               "near_edge": (np.ctypeslib.as_array(res.near_edge, (res.n_near,)).copy() if res.n_near
                             else np.zeros(0, np.uint64)),
               "near_score64": (np.ctypeslib.as_array(res.near_score64, (res.n_near,)).copy() if res.n_near
                                else np.zeros(0, np.float64))}
        for k, dt in (("from", np.uint32), ("to", np.uint32), ("scomp", np.float32)):
            if out[k] is None:
                out[k] = np.zeros(0, dt)
        self._L.rbg_graph_result_free(C.byref(res))
        return out, d

    def release(self, session):
        st = self._L.rbg_server_release(self._h, session)
        if st != 0:
            raise RbgError(f"rbg_server_release: status {st}: {self._err()}")

    def stats(self):
        s = ServerStats()
        st = self._L.rbg_server_get_stats(self._h, C.byref(s))
        if st != 0:
            raise RbgError(f"rbg_server_get_stats: status {st}")
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def stop(self):
        if self._h:
            self._L.rbg_server_stop(self._h)
            self._h = C.c_void_p()
            self._inputs.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


def depth_features(depth, min_cv=1.0):
    """rbg_depth_features: ranked, unit, raw, colnorm, anynz."""
    d = np.ascontiguousarray(depth, np.float32)
    n, S = d.shape
    ranked = np.zeros_like(d)
    unit = np.zeros_like(d)
    raw = np.zeros_like(d)
    colnorm = np.zeros(S, np.float64)
    anynz = np.zeros(n, np.uint8)
    st = lib().rbg_depth_features(_p(d, f32p), n, S, min_cv, _p(ranked, f32p), _p(unit, f32p), _p(raw, f32p),
                                  _p(colnorm, f64p), _p(anynz, u8p))
    if st != 0:
        raise RbgError(f"rbg_depth_features: status {st}")
    return ranked, unit, raw, colnorm, anynz


N_CANON4 = 136  # canonical 4-mers (min of a code and its reverse complement)


def sketch_tables(m=500, seed=42):
    """rbg_sketch_tables: (T [136, m] f64, W [136, m] u64, cls [m, 136] u8).

    Element x is the x-th canonical 4-mer in ascending code order; T/W are its
    proposal values at weight 1 and its winner identities (src/probmh.cpp:634-686)."""
    T = np.zeros((N_CANON4, m), np.float64)
    W = np.zeros((N_CANON4, m), np.uint64)
    cls = np.zeros((m, N_CANON4), np.uint8)
    st = lib().rbg_sketch_tables(m, seed, _p(T, f64p), _p(W, u64p), _p(cls, u8p))
    if st != 0:
        raise RbgError(f"rbg_sketch_tables(m={m}): status {st}")
    return T, W, cls


def calibrate(rowmax_cnt, m, b0, regime, coverage=0.95, rand_seed=1):
    """rbg_calibrate: simCutoff (x1000) from per-row maximum match counts."""
    rm = np.ascontiguousarray(rowmax_cnt, np.uint32)
    out = C.c_double()
    st = lib().rbg_calibrate(_p(rm, u32p), len(rm), m, b0, coverage, rand_seed, regime, C.byref(out))
    if st != 0:
        raise RbgError(f"rbg_calibrate(n={len(rm)}, regime={regime}): status {st}")
    return out.value
