"""Pins of the CPU oracle (oracle/) and of the product's host-side tables.

Every check compares against something the oracle did not compute itself:
  * the reference's own probmh.cpp (tests/golden/probmh_golden.npz, written
    by tests/golden/make_golden.py from oracle/_ref; and oracle/_ref live
    when it is built),
  * glibc's rand() and the C++ standard's mt19937_64 known answer,
  * AVX-512 intrinsics compiled by the host compiler (the reference's gate),
  * brute force in numpy / Python on small inputs, closed forms and invariants.
No GPU is needed.
"""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
G = np.load(os.path.join(HERE, "golden", "probmh_golden.npz"))
needs_ref = pytest.mark.skipif(not orc.ref_available(), reason="oracle/_ref not built (reference absent)")

ENC = {c: i for i, c in enumerate("ACGT")}
ENC.update({c.lower(): i for c, i in list(ENC.items())})
M64 = (1 << 64) - 1


def canonical_codes():
    out = []
    for v in range(256):
        rc = 0
        for d in range(4):
            rc = (rc << 2) | (3 - ((v >> (2 * d)) & 3))
        if v <= rc:
            out.append(v)
    return np.array(out, np.uint64)


CODES = canonical_codes()


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


# ---------------------------------------------------------------------------
# RNG primitives (src/probmh.cpp:60-87, 319-357; src/hash_int.h:79-87)
# ---------------------------------------------------------------------------

def test_canonical_code_count():
    assert len(CODES) == 136


def test_fmix64_golden():
    got = [orc.fmix64(int(a), int(b)) for a, b in G["fmix_in"]]
    assert np.array_equal(np.array(got, np.uint64), G["fmix_out"])


def test_wyrand_and_uniform_int_golden():
    for si, s in enumerate(G["wy_states"]):
        out, _ = orc.wyrand_seq(int(s), 32)
        assert np.array_equal(out, G["wyrand"][si])
        for bi, b in enumerate(G["uni_bounds"]):
            out, _ = orc.wy_uniform_int_seq(int(s), int(b), 32)
            assert np.array_equal(out, G["uni"][si, bi]), (s, b)
            assert (out < b).all()


def test_zig_and_ted_golden():
    for t, s in enumerate((7, 0xC0FFEE)):
        out, _ = orc.zig_seq(s, 512)
        assert np.array_equal(bits(out), bits(G["zig"][t]))
    rows = [0, 1, 10, 100, 497]
    for t, r in enumerate(rows):
        c1, c2, c3 = G["ted_c"][t]
        out, _ = orc.ted_seq(c1, c2, c3, 1234 + r, 64)
        assert np.array_equal(bits(out), bits(G["ted"][t]))


def _header_table(name):
    src = open(os.path.join(REPO, "rabbitbin_amd", "csrc", "zig_tables.h")).read()
    body = re.search(name + r"\[257\]\s*=\s*\{(.*?)\};", src, re.S).group(1)
    vals = [float.fromhex(t.strip()) for t in body.split(",") if t.strip()]
    return np.array(vals, np.float64)


def test_generated_ziggurat_tables_match_reference():
    """rabbitbin_amd/csrc/zig_tables.h is generated from the Marsaglia-Tsang
    construction; it must equal the reference's tables bit for bit."""
    assert np.array_equal(bits(_header_table("kZigX")), bits(G["zig_x"]))
    assert np.array_equal(bits(_header_table("kZigY")), bits(G["zig_y"]))


@needs_ref
def test_rng_streams_vs_reference_random_states():
    rng = np.random.default_rng(5)
    ted, _ = orc.ted_params(500)
    ted16, _ = orc.ted_params(16)
    for _ in range(40):
        s = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        a, sa = orc.wyrand_seq(s, 64)
        b, sb = orc.wyrand_seq(s, 64, "ref")
        assert np.array_equal(a, b) and sa == sb
        bound = int(rng.integers(1, 2**32))
        a, sa = orc.wy_uniform_int_seq(s, bound, 64)
        b, sb = orc.wy_uniform_int_seq(s, bound, 64, "ref")
        assert np.array_equal(a, b) and sa == sb
        a, sa = orc.zig_seq(s, 2000)  # ~1% of draws take the slow (exp) path
        b, sb = orc.zig_seq(s, 2000, "ref")
        assert np.array_equal(bits(a), bits(b)) and sa == sb
        for tab in (ted, ted16):
            r = int(rng.integers(0, len(tab)))
            a, sa = orc.ted_seq(*tab[r, 2:5], s, 64)
            b, sb = orc.ted_seq(*tab[r, 2:5], s, 64, "ref")
            assert np.array_equal(bits(a), bits(b)) and sa == sb
        x = int(rng.integers(0, 2**63))
        assert orc.fmix64(x, 42) == orc.fmix64(x, 42, "ref")


# ---------------------------------------------------------------------------
# Sketches (A1-A3, A6)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("m", [2, 16, 500])
def test_streaming_sketch_golden(m):
    """Oracle's streaming ProbMinHash4 == the reference's, bit for bit."""
    W = G[f"sk_m{m}_w"]
    for t in range(len(W)):
        sel = W[t] > 0
        win, reg = orc.pmh_stream(CODES[sel], W[t][sel], m)
        assert np.array_equal(win, G[f"sk_m{m}_win"][t]), t
        assert np.array_equal(bits(reg), bits(G[f"sk_m{m}_reg"][t])), t


@pytest.mark.parametrize("m", [2, 16, 500])
def test_product_sketch_tables_golden(m):
    """rbg_sketch_tables (host code of the product) == the reference's
    single-element sketches."""
    from rabbitbin_amd import api
    T, W, _ = api.sketch_tables(m)
    if m == 500:
        assert hashlib.sha256(T.tobytes()).digest() == G["tw_m500_sha256_T"].tobytes()
        assert hashlib.sha256(W.tobytes()).digest() == G["tw_m500_sha256_W"].tobytes()
    else:
        assert np.array_equal(bits(T), bits(G[f"tw_m{m}_T"]))
        assert np.array_equal(W, G[f"tw_m{m}_W"])


def fold32(h):
    h = np.asarray(h, np.uint64)
    return (h ^ (h >> np.uint64(32))).astype(np.uint32)


def fold16_def(w):
    """fold_win16 (src/rabbitbin.cpp:249-253)."""
    w = np.asarray(w, np.uint32)
    h = (w ^ (w >> np.uint32(16))).astype(np.uint16)
    h = np.where(h == 0, np.uint16(1), h)
    return np.where(w == 0, np.uint16(0), h).astype(np.uint16)


def test_fold16():
    rng = np.random.default_rng(3)
    w = rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32)
    w[:4] = [0, 0x12341234, 1, 0xFFFF0000]
    assert np.array_equal(orc.fold16(w), fold16_def(w))
    assert list(fold16_def(w[:4])) == [0, 1, 1, 0xFFFF]


@pytest.mark.parametrize("m", [2, 16, 500])
def test_class_bytes_definition(m):
    """cls[k][x] = smallest y with fold16(W[y][k]) == fold16(W[x][k]); 0xFF
    when that fold is 0.  Byte equality of classes (excluding 0xFF) must equal
    the reference's u16 test a == b && a != 0, for all (k, x, y)."""
    from rabbitbin_amd import api
    _, W, cls = api.sketch_tables(m)
    f = fold16_def(fold32(W)).T  # [m, 136]
    for k in range(m):
        eq16 = (f[k][:, None] == f[k][None, :]) & (f[k][:, None] != 0)
        c = cls[k]
        eq8 = (c[:, None] == c[None, :]) & (c[:, None] != 0xFF)
        assert np.array_equal(eq16, eq8), k
        for x in range(136):
            if f[k][x] == 0:
                assert c[x] == 0xFF
            else:
                assert c[x] == np.flatnonzero(f[k] == f[k][x])[0]


@pytest.mark.parametrize("m", [16, 500])
def test_closed_form_sketch_equals_reference_stream(m):
    """A3': leaf[k] = min_x (1/w_x) * T[x][k] (ties to the smaller x) and
    winner[k] = W[x*][k] reproduces the reference's streaming sketch.  This is
    the algorithm of the GPU sketch kernel."""
    from rabbitbin_amd import api
    T, W, _ = api.sketch_tables(m)
    WT = G[f"sk_m{m}_w"]
    for t in range(len(WT)):
        w = WT[t]
        sel = np.flatnonzero(w > 0)
        hv = (1.0 / w[sel])[:, None] * T[sel]
        arg = np.argmin(hv, axis=0)  # first minimum = smallest element
        leaf = hv[arg, np.arange(m)]
        win = W[sel[arg], np.arange(m)]
        assert np.array_equal(win, G[f"sk_m{m}_win"][t]), t
        assert np.array_equal(bits(leaf), bits(G[f"sk_m{m}_reg"][t])), t


def kmer4_brute(seq):
    """Canonical 4-mer histogram by definition (src/rabbitbin.cpp:1459-1484):
    windows of 4 valid bases (ACGT, either case), code = bases MSB first,
    canonical = min(code, reverse-complement code)."""
    cnt = np.zeros(256, np.uint32)
    base = np.zeros(4, np.uint64)
    tot = 0
    s = seq.decode("latin-1")
    for c in s:
        if c in ENC:
            base[ENC[c]] += 1
    for i in range(len(s) - 3):
        w = s[i:i + 4]
        if all(c in ENC for c in w):
            d = [ENC[c] for c in w]
            fwd = sum(b << (2 * (3 - t)) for t, b in enumerate(d))
            rc = [3 - b for b in reversed(d)]
            rev = sum(b << (2 * (3 - t)) for t, b in enumerate(rc))
            cnt[min(fwd, rev)] += 1
            tot += 1
    return cnt, base, tot


def random_seq(rng, L, n_frac=0.0, lower=0.0, junk=0.0):
    s = np.array(list("ACGT"))[rng.integers(0, 4, L)]
    if lower:
        m = rng.random(L) < lower
        s[m] = np.char.lower(s[m])
    if n_frac:
        s[rng.random(L) < n_frac] = "N"
    if junk:
        s[rng.random(L) < junk] = rng.choice(list("RYKM-.*xn"))
    return "".join(s).encode()


SEQ_CASES = [b"", b"A", b"ACG", b"ACGT", b"NNNN", b"acgt", b"ACGTACGT", b"AAAATTTT", b"ACNGTACGTNNACGTA",
             b"GGGGCCCCGGGG", b"ACGT" * 50, b"XYZACGTZ"]


def test_kmer4_hist_bruteforce():
    rng = np.random.default_rng(11)
    cases = SEQ_CASES + [random_seq(rng, int(rng.integers(0, 600)), n_frac=f, lower=lo, junk=j)
                         for f, lo, j in [(0, 0, 0), (0.05, 0, 0), (0, 0.5, 0), (0.02, 0.2, 0.02)] * 5]
    canon = np.zeros(256, bool)
    canon[CODES.astype(int)] = True
    for s in cases:
        cnt, base, tot = orc.kmer4_hist(s)
        bc, bb, bt = kmer4_brute(s)
        assert np.array_equal(cnt, bc), s[:40]
        assert np.array_equal(base, bb) and tot == bt
        assert not cnt[~canon].any()


def weights_closed_form(cnt, base, total, cap=20.0):
    """GC-normalised weights, mode 1 (src/rabbitbin.cpp:1491-1516), in the
    reference's operation order (digits LSB first, products from 1.0)."""
    vb = int(base.sum())
    p = [float(base[b]) / float(vb) if vb > 0 else 0.25 for b in range(4)]
    inv = 1.0 / float(total)
    w = np.zeros(256)
    for v in range(256):
        if not cnt[v]:
            continue
        pf = pr = 1.0
        code = v
        for _ in range(4):
            b = code & 3
            pf *= p[b]
            pr *= p[3 - b]
            code >>= 2
        obs = float(cnt[v]) * inv
        ex = pf + pr
        w[v] = min(obs / ex, cap) if ex > 1e-14 else 0.0
    return w


def test_kmer4_weights_closed_form():
    rng = np.random.default_rng(12)
    for s in SEQ_CASES[3:] + [random_seq(rng, int(rng.integers(4, 3000)), n_frac=0.01) for _ in range(30)] + \
            [b"A" * 100 + b"C", b"AAAAAAAAAT", b"GCGCGCGCGCGCGCA"]:
        cnt, base, tot = kmer4_brute(s)
        if not tot:
            continue
        assert np.array_equal(bits(orc.kmer4_weights(cnt, base, tot)), bits(weights_closed_form(cnt, base, tot)))


@needs_ref
def test_build_pmh_winners_vs_reference_stream():
    """Whole stage (a) for one contig: brute-force histogram, closed-form
    weights, the reference's own sketch, 64->32 fold == the oracle."""
    rng = np.random.default_rng(13)
    seqs = [random_seq(rng, int(L), n_frac=0.003, lower=0.1) for L in rng.integers(4, 6000, 12)]
    seqs += [b"ACGTNACGT", b"AAAAAAAAAAAA", b"acgtacgtac"]
    for s in seqs:
        cnt, base, tot = kmer4_brute(s)
        w = weights_closed_form(cnt, base, tot) if tot else np.zeros(256)
        present = np.flatnonzero(w > 0)
        win64, _ = orc.pmh_stream(present.astype(np.uint64), w[present], 500, which="ref")
        out = np.zeros(500, np.uint32)
        orc.orc().orc_build_pmh_winners(s, len(s), 500, 42, out.ctypes.data_as(orc.u32p))
        assert np.array_equal(out, fold32(win64)), s[:30]


# ---------------------------------------------------------------------------
# glibc rand (A11) and libstdc++ mt19937_64 + uniform_int_distribution (A7)
# ---------------------------------------------------------------------------

LIBC = ctypes.CDLL("libc.so.6")


def test_rand_known_answer():
    assert list(orc.rand_seq(1, 5)) == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]


def test_rand_stream_equals_libc():
    for seed in (1, 2, 42, 12345, 2**31 - 1):
        LIBC.srand(seed)
        want = [LIBC.rand() for _ in range(3000)]
        assert list(orc.rand_seq(seed, 3000)) == want


def mt19937_64(seed=5489):
    """MT19937-64 (Matsumoto & Nishimura 2004; the C++ std::mt19937_64)."""
    NN, MM, A = 312, 156, 0xB5026F5AA96619E9
    UM, LM = 0xFFFFFFFF80000000, 0x7FFFFFFF
    mt = [seed & M64]
    for i in range(1, NN):
        mt.append((6364136223846793005 * (mt[-1] ^ (mt[-1] >> 62)) + i) & M64)
    idx = NN
    while True:
        if idx >= NN:
            for i in range(NN):
                x = (mt[i] & UM) | (mt[(i + 1) % NN] & LM)
                xa = (x >> 1) ^ (A if x & 1 else 0)
                mt[i] = mt[(i + MM) % NN] ^ xa
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        yield y & M64


def uniform_int(gen, n):
    """uniform_int_distribution<size_t>(0, n-1) over a 64-bit engine in
    libstdc++ (GCC 11): Lemire's nearly divisionless method with a 128-bit
    product (D. Lemire, "Fast Random Integer Generation in an Interval", 2019)."""
    prod = next(gen) * n
    low = prod & M64
    if low < n:
        thr = ((1 << 64) - n) % n
        while low < thr:
            prod = next(gen) * n
            low = prod & M64
    return prod >> 64


def test_mt19937_64_known_answer():
    g = mt19937_64()
    for _ in range(9999):
        next(g)
    assert next(g) == 9981545732273789042  # C++ standard [rand.predef]
    assert orc.mt19937_64_nth(10000) == 9981545732273789042


@pytest.mark.parametrize("n", [3, 2000, 50000, 300000, 2**33 + 5])
def test_baseline_draws(n):
    g = mt19937_64(0x9E3779B97F4A7C15)
    want_i, want_j = [], []
    for _ in range(30000):
        want_i.append(uniform_int(g, n))
        want_j.append(uniform_int(g, n))
    i, j = orc.baseline_draws(n)
    assert list(i) == want_i and list(j) == want_j


# ---------------------------------------------------------------------------
# Similarity (A8), gate dot product (A9)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("b0", [0.0, 0.02, 0.31, 0.9])
def test_sv_from_count_properties(b0):
    m = 500
    sv = np.array([orc.sv_from_count(c, m, b0) for c in range(m + 1)])
    assert (np.diff(sv) >= 0).all()
    assert sv[m] == 1 - 1e-6
    assert (sv[np.arange(m + 1) * (1.0 / m) <= b0] == 0).all()
    assert (sv[np.arange(m + 1) * (1.0 / m) > b0] > 0).all()
    if b0 == 0.0:
        assert np.array_equal(sv[:m], np.arange(m) * (1.0 / m))


def test_match_count_bruteforce():
    rng = np.random.default_rng(4)
    for m in (2, 16, 500):
        a = rng.integers(0, 4, m).astype(np.uint16)
        b = rng.integers(0, 4, m).astype(np.uint16)
        want = int(((a == b) & (a != 0)).sum())
        got = orc.orc().orc_match_count16(a.ctypes.data_as(orc.u16p), b.ctypes.data_as(orc.u16p), m)
        assert got == want


def test_unit_dot_emulation_equals_avx512():
    """The oracle's scalar restatement of the reference's AVX-512 unit_dot_f
    (two 16-lane FMA accumulators, masked tail, GCC's reduce) == the same
    loop written with AVX-512 intrinsics, bit for bit, for every S <= 64."""
    rng = np.random.default_rng(6)
    if orc.unit_dot_f_avx512(np.ones(3, np.float32), np.ones(3, np.float32)) is None:
        pytest.skip("host lacks AVX-512F")
    for S in range(1, 65):
        for _ in range(60):
            a = rng.normal(size=S).astype(np.float32)
            b = (a * rng.uniform(0.5, 1.5) + rng.normal(scale=0.3, size=S)).astype(np.float32)
            a /= np.float32(np.linalg.norm(a))
            b /= np.float32(np.linalg.norm(b))
            e = orc.unit_dot_f(a, b)
            v = orc.unit_dot_f_avx512(a, b)
            assert e.view(np.uint32) == v.view(np.uint32), (S, e, v)
            assert abs(float(e) - float(np.dot(a.astype(np.float64), b))) < 1e-5


def test_depth_corr_emulation_equals_avx512():
    """The oracle's edge-weight correlation in the reference order
    (depth_corr_fast: FP32 chunks of 32/16 reduced once, FP64 tail) == the
    same loop with AVX-512 intrinsics, bit for bit, for S = 1..70; below 16 it
    is the FP64 sum."""
    rng = np.random.default_rng(16)
    if orc.depth_corr_avx512(np.ones(3, np.float32), np.ones(3, np.float32)) is None:
        pytest.skip("host lacks AVX-512F")
    for S in range(1, 71):
        for _ in range(60):
            a = rng.normal(size=S).astype(np.float32)
            b = (a * rng.uniform(-1.5, 1.5) + rng.normal(scale=0.5, size=S)).astype(np.float32)
            a /= np.float32(np.linalg.norm(a))
            b /= np.float32(np.linalg.norm(b))
            e = orc.depth_corr(a, b, 0)
            v = orc.depth_corr_avx512(a, b)
            assert np.float64(e).view(np.uint64) == np.float64(v).view(np.uint64), (S, e, v)
            if S < 16:
                assert e == orc.depth_corr(a, b, 1)
            assert abs(e - float(np.dot(a.astype(np.float64), b))) < 1e-5


# ---------------------------------------------------------------------------
# Calibration (A11)
# ---------------------------------------------------------------------------

def py_converge(maxsim, coverage, rand):
    """calib_converge (src/impl/rb_graph.cpp:390-415)."""
    maxsim = np.asarray(maxsim, np.float64)
    p, pp, pcov = 999, 1000, 0.0
    while p > 50:
        cov = np.count_nonzero(maxsim >= p / 1000.0) / len(maxsim)
        if cov >= coverage:
            if cov - coverage > coverage - pcov:
                p = pp
            break
        pp, pcov = p, cov
        r = rand() % 3
        p -= r + 1 if p > 990 else (r + 3 if p > 900 else r + 9)
    return p


def py_random_unique(idx, num, rand):
    left = len(idx)
    for s in range(num):
        r = rand() % left
        idx[s], idx[s + r] = idx[s + r], idx[s]
        left -= 1


def py_calibrate(rowmax, m, b0, regime, coverage, seed):
    """calibrate_sim_cutoff(coverage, true) (regime 1, src/impl/rb_graph.cpp:
    1362-1384) and the sampled calibration of gen_fused_calib_graph (regime 2,
    :502-525, 1243-1254), driven by glibc's own rand()."""
    LIBC.srand(seed)
    rand = LIBC.rand
    n = len(rowmax)
    svd = np.array([orc.sv_from_count(c, m, b0) for c in range(m + 1)])
    if regime == 1:
        idx = list(range(n))
        py_random_unique(idx, n, rand)
        return float(py_converge([svd[rowmax[i]] if n > 1 else 0.0 for i in idx], coverage, rand))
    rounds = []
    for _ in range(10):
        idx = list(range(n))
        py_random_unique(idx, 2500, rand)
        rounds.append(idx[:2500])
    round_of = {}
    for r in range(9, -1, -1):
        for li, c in enumerate(rounds[r]):
            round_of[c] = (r, li)
    maxsim = np.zeros((10, 2500), np.float32)
    for c, (r, li) in round_of.items():
        maxsim[r, li] = np.float32(svd[rowmax[c]])
    sum_p = 0.0
    for r in range(10):
        p = float(py_converge(maxsim[r].astype(np.float64), coverage, rand))
        if p < 51.0:
            p = 50.0
        sum_p += p
        if r == 1 and sum_p / 2 < 51.0:
            sum_p = 500.0
            break
    return sum_p / 10.0


def rowmax_cases(rng, n, m):
    yield rng.integers(0, m + 1, n).astype(np.uint32)
    yield np.minimum(m, rng.binomial(m, 0.5, n)).astype(np.uint32)
    yield np.where(rng.random(n) < 0.97, rng.integers(int(0.9 * m), m + 1, n), 0).astype(np.uint32)
    yield np.full(n, m, np.uint32)
    yield np.zeros(n, np.uint32)


def test_calibrate_small_restated_with_libc():
    from rabbitbin_amd import api
    rng = np.random.default_rng(7)
    for n in (1, 2, 50, 800):
        for rm in rowmax_cases(rng, n, 500):
            for seed, b0 in ((1, 0.0), (9, 0.05)):
                want = py_calibrate(rm, 500, b0, 1, 0.95, seed)
                assert orc.calibrate(rm, 500, b0, 1, 0.95, seed) == want
                assert api.calibrate(rm, 500, b0, 1, 0.95, seed) == want


def test_calibrate_fused_restated_with_libc():
    from rabbitbin_amd import api
    rng = np.random.default_rng(8)
    for n in (2500, 4000):
        for rm in rowmax_cases(rng, n, 500):
            want = py_calibrate(rm, 500, 0.03, 2, 0.95, 1)
            assert orc.calibrate(rm, 500, 0.03, 2, 0.95, 1) == want
            assert api.calibrate(rm, 500, 0.03, 2, 0.95, 1) == want


def test_calibrate_closed_forms():
    from rabbitbin_amd import api
    full = np.full(3000, 500, np.uint32)
    zero = np.zeros(3000, np.uint32)
    assert orc.calibrate(full, 500, 0.0, 1) == 999.0
    assert api.calibrate(full, 500, 0.0, 1) == 999.0
    # Fused: a contig drawn in several rounds counts only in its lowest round
    # (its other slots stay 0).  At n = 1e6 fewer than 2.3% of any round's
    # slots are such repeats, so every round reaches 95% at p = 999 ...
    big = np.full(1_000_000, 500, np.uint32)
    assert orc.calibrate(big, 500, 0.0, 2) == 999.0
    assert api.calibrate(big, 500, 0.0, 2) == 999.0
    # ... while at n = 3000 later rounds are mostly repeats and stay below it.
    assert api.calibrate(full, 500, 0.0, 2) < 999.0
    assert api.calibrate(zero, 500, 0.0, 2) == 50.0  # floor, and the early stop after round 2
    p = api.calibrate(zero, 500, 0.0, 1)
    assert 40 < p <= 50
    with pytest.raises(api.RbgError):
        api.calibrate(np.zeros(100, np.uint32), 500, 0.0, 2)  # 10 x 2500 samples need n >= 2500
    with pytest.raises(api.RbgError):
        api.calibrate(np.full(10, 501, np.uint32), 500, 0.0, 1)


# ---------------------------------------------------------------------------
# Depth features
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("S", [1, 2, 3, 10, 30])
def test_depth_features(S):
    from scipy.stats import rankdata
    from rabbitbin_amd import api
    rng = np.random.default_rng(S)
    n = 400
    d = rng.poisson(rng.lognormal(1, 1, (n, 1)) * rng.lognormal(0, 1, (1, S)) * 3).astype(np.float32)
    d[::7] = np.round(d[::7] / 4)  # many ties
    d[5] = 0.0
    d[6] = 2.5
    o = orc.depth_features(d)
    p = api.depth_features(d)
    for a, b in zip(o, p):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    ranked, unit, raw, colnorm, anynz = o
    if S == 1:
        assert np.array_equal(ranked, d) and not unit.any()
        assert np.array_equal(anynz, (d[:, 0] > 1.0).astype(np.uint8))
        return
    assert np.array_equal(ranked, rankdata(d, method="average", axis=1).astype(np.float32))
    assert anynz.all()  # ranks > 1.0 exist in every row once S >= 2
    c = ranked.astype(np.float64) - ranked.mean(axis=1, keepdims=True)
    nrm = np.linalg.norm(c, axis=1, keepdims=True)
    want = np.divide(c, nrm, out=np.zeros_like(c), where=nrm > 0)
    assert np.abs(unit - want).max() < 1e-6
    assert np.array_equal(raw, d)
    mean = d.astype(np.float64).mean(axis=0)
    assert np.allclose(colnorm, np.where(mean > 1e-12, 1.0 / mean, 0.0), rtol=1e-12)


# ---------------------------------------------------------------------------
# Graph construction (A10, A12) by brute force on small inputs
# ---------------------------------------------------------------------------

def planted_win16(rng, n, m, groups):
    centre = rng.integers(1, 2**16, (groups, m)).astype(np.uint16)
    g = rng.integers(0, groups, n)
    keep = rng.random((n, 1)) * 0.9 + 0.05
    w = np.where(rng.random((n, m)) < keep, centre[g], rng.integers(1, 2**16, (n, m))).astype(np.uint16)
    w[rng.random((n, m)) < 0.02] = 0
    return w, g


def counts_matrix(w):
    n, m = w.shape
    out = np.zeros((n, n), np.int32)
    for i0 in range(0, n, 128):
        a = w[i0:i0 + 128]
        out[i0:i0 + 128] = ((a[:, None, :] == w[None, :, :]) & (a[:, None, :] != 0)).sum(2)
    return out


def gate_dot_s3(u):
    """unit_dot_f for S = 3 (one masked 16-lane FMA into zero, then GCC's
    reduce: ((p0 + p2) + p1) with p = fl(a*b))."""
    p = (u[:, None, :] * u[None, :, :]).astype(np.float32)
    return (p[..., 0] + p[..., 2]) + p[..., 1]


def topk_rows(sv, valid, K):
    n = sv.shape[0]
    top = []
    for i in range(n):
        js = np.flatnonzero(valid[i])
        order = np.lexsort((js, -sv[i, js].astype(np.float64)))  # sv desc, id asc
        top.append(js[order[:K]])
    return top


def check_graph_bruteforce(w16, unit, S, regime, K):
    n, m = w16.shape
    b0 = orc.pmh_baseline(w16)
    g = orc.build_graph(w16, unit, None if unit is None else np.ones(n, np.uint8), S, b0, K=K,
                        force_regime=regime)
    cnt = counts_matrix(w16)
    svf = np.array([orc.sv_from_count(c, m, b0) for c in range(m + 1)], np.float32)
    sv = svf[cnt]
    valid = ~np.eye(n, dtype=bool)
    if regime == 2:
        dot = gate_dot_s3(unit)
        valid &= ~(dot < np.float32(0.7 - 1e-4))
    rowmax = np.where(valid, cnt, 0).max(1)
    assert np.array_equal(g["rowmax_cnt"], rowmax)
    want_cut = py_calibrate(rowmax, m, b0, regime, 0.95, 1)
    assert g["sim_cutoff"] == want_cut
    top = topk_rows(sv, valid, K)
    for i in range(n):
        L = g["topk_len"][i]
        assert L == len(top[i])
        assert set(g["topk_id"][i, :L]) == set(top[i])
    cut = want_cut / 1000.0
    kept = [set(j for j in top[i] if float(sv[i, j]) > cut) for i in range(n)]
    if regime == 1:
        want = {(i, j) for i in range(n) for j in kept[i] if j > i}
    else:
        want = {(i, j) for i in range(n) for j in kept[i] if j > i and i in kept[j]}
    got = set(zip(g["from"].tolist(), g["to"].tolist()))
    assert got == want
    for f, t, s in zip(g["from"], g["to"], g["sv"]):
        assert s == sv[f, t]
    return g


def test_small_regime_bruteforce():
    rng = np.random.default_rng(21)
    w, _ = planted_win16(rng, 600, 64, 12)
    g = check_graph_bruteforce(w, None, 0, 1, K=9)
    assert len(g["from"]) > 0


def test_fused_regime_bruteforce():
    rng = np.random.default_rng(22)
    n = 2600
    w, grp = planted_win16(rng, n, 32, 40)
    ab = rng.lognormal(0, 1.5, (40, 3))
    d = (ab[grp] * rng.lognormal(0, 0.3, (n, 3))).astype(np.float32)
    _, unit, _, _, _ = orc.depth_features(d)
    g = check_graph_bruteforce(w, unit, 3, 2, K=7)
    assert len(g["from"]) > 0 and 0 < g["gated_pairs"] < n * (n - 1) // 2
    # the per-row restatement used for sampled rows of large assemblies
    rows = np.array([0, 5, 100, 1300, n - 1], np.uint32)
    ids, ln, rm, gated = orc.rows_topk(w, unit, np.ones(n, np.uint8), 3, orc.pmh_baseline(w), rows, K=7)
    for t, r in enumerate(rows):
        assert ln[t] == g["topk_len"][r] and rm[t] == g["rowmax_cnt"][r]
        assert set(ids[t, :ln[t]].tolist()) == set(g["topk_id"][r, :ln[t]].tolist())
    ok = ~(gate_dot_s3(unit) < np.float32(0.7 - 1e-4))
    np.fill_diagonal(ok, False)
    assert np.array_equal(gated, ok[rows].sum(1))


# ---------------------------------------------------------------------------
# Edge weights (A13)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("S", [2, 3, 10, 20, 30])
def test_edge_weights_closed_form(S):
    """FP64 mode == Pearson on average ranks, min with the weighted Jaccard of
    column-normalised depths, -0.3 cut, 0.70 floor; reference arithmetic
    (mode 0) within 1e-6 of it."""
    from scipy.stats import rankdata
    rng = np.random.default_rng(100 + S)
    n = 300
    base = rng.lognormal(1, 1, (30, S))
    grp = rng.integers(0, 30, n)
    d = (base[grp] * rng.lognormal(0, 0.25, (n, S))).astype(np.float32)
    ranked, unit, raw, colnorm, _ = orc.depth_features(d)
    ne = 4000
    fr = rng.integers(0, n, ne).astype(np.uint32)
    to = rng.integers(0, n, ne).astype(np.uint32)
    sc = rng.random(ne).astype(np.float32)
    w0, _ = orc.edge_weights(fr, to, sc, unit, raw, colnorm, S, mode=0)
    w1, w64 = orc.edge_weights(fr, to, sc, unit, raw, colnorm, S, mode=1)
    rk = rankdata(d, method="average", axis=1)
    want = np.zeros(ne)
    for e in range(ne):
        a, b = rk[fr[e]], rk[to[e]]
        ca, cb = a - a.mean(), b - b.mean()
        den = np.sqrt((ca * ca).sum() * (cb * cb).sum())
        corr = (ca * cb).sum() / den if den > 0 else 0.0
        if corr < -0.3:
            continue
        x = raw[fr[e]] * colnorm
        y = raw[to[e]] * colnorm
        wj = np.minimum(x, y).sum() / np.maximum(x, y).sum()
        # auto composition weight: 1 for S <= 2 (the edge keeps sComp), else 0
        v = float(sc[e]) if S <= 2 else min(max(corr, 0.0), min(max(wj, 0.0), 1.0))
        want[e] = v if v >= 0.70 else 0.0
    near = np.abs(want - 0.70) < 1e-6
    assert np.abs(w64 - want)[~near].max() < 1e-6
    assert np.abs(w0.astype(np.float64) - w64)[~near].max() < 1e-6
    if S > 2:
        assert (w64 == 0).any() and (w64 >= 0.70).any()
