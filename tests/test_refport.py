"""The performance port of the reference's fused pass (oracle/rbg_refport_fused.cpp,
after src/impl/rb_graph.cpp:501-1356), which bench.py times as the CPU
baseline: its graph must equal the oracle's row-wise fused regime, edge for
edge, with the same simCutoff and evaluated-pair count, in both its builds
(AVX-512 with the int8 VNNI prefilter, and scalar) and on the paths its
shortcuts take: the int8 prefilter (S >= 16), the composition early exit
(full heaps: small K), no gate (S = 0, and S = 2 where w_comp = 1), and
rows flagged all-zero."""
import numpy as np
import pytest

from oracle import oracle as orc
from rabbitbin_amd import synth

BUILDS = [False] + ([True] if orc.host_avx512_vnni() else [])


def _inputs(n, S, seed):
    seqs, off, depth, _ = synth.make_assembly(n, 30, np.log(8000), 0.8, max(S, 1), seed=seed)
    w16 = orc.fold16(orc.sketch_batch(seqs, off))
    b0 = orc.pmh_baseline(w16)
    if S == 0:
        return w16, None, None, b0
    depth = depth[:, :S].copy()
    _, unit, _, _, anynz = orc.depth_features(depth)
    return w16, unit, anynz, b0


def _same(w16, unit, anynz, S, b0, K, avx):
    want = orc.build_graph(w16, unit, anynz, S, b0, K=K, force_regime=2)
    got = orc.port_fused_graph(w16, unit, anynz, S, b0, K=K, avx512=avx)
    o = np.lexsort((got["to"], got["from"]))
    assert got["sim_cutoff"] == want["sim_cutoff"]
    assert got["gated_pairs"] == want["gated_pairs"]
    assert np.array_equal(got["from"][o], want["from"])
    assert np.array_equal(got["to"][o], want["to"])
    assert np.array_equal(got["sv"][o].view(np.uint32), want["sv"].view(np.uint32))
    assert got["stats"]["avx512"] == int(avx)
    return got["stats"]


@pytest.mark.parametrize("avx", BUILDS)
@pytest.mark.parametrize("S", [10, 30])
def test_port_equals_oracle(S, avx):
    w16, unit, anynz, b0 = _inputs(3000, S, 11 + S)
    st = _same(w16, unit, anynz, S, b0, 200, avx)
    assert st["q8"] == int(avx and S >= 16)
    if st["q8"]:
        assert st["q8_pruned"] > 0


@pytest.mark.parametrize("avx", BUILDS)
def test_port_early_exit_with_full_heaps(avx):
    """Registers drawn at random with planted groups of 6 identical rows and
    K = 4: grouped rows fill their heaps with similarity 1 - 1e-6 (all ties,
    ordered by id through the packed key), after which the 64-register bound
    stops their pairs with unrelated rows.  12,000 rows leave about a tenth
    outside the calibration samples, where the early exit applies."""
    rng = np.random.default_rng(7)
    n, m = 12000, 500
    w16 = rng.integers(1, 65536, (n, m), dtype=np.uint16)
    groups = rng.permutation(n)[:6 * 1500].reshape(1500, 6)
    w16[groups[:, 1:]] = w16[groups[:, :1]]
    b0 = orc.pmh_baseline(w16)
    st = _same(w16, None, None, 0, b0, 4, avx)
    assert st["ub_pruned"] > 0


@pytest.mark.parametrize("avx", BUILDS)
@pytest.mark.parametrize("S", [0, 2])
def test_port_without_gate(S, avx):
    w16, unit, anynz, b0 = _inputs(2600, S, 21 + S)
    st = _same(w16, unit, anynz, S, b0, 30, avx)
    assert st["float_checked"] == 0 and st["q8"] == 0


@pytest.mark.parametrize("avx", BUILDS)
def test_port_zero_rows(avx):
    """Rows flagged all-zero pair only with flagged-non-zero rows (is_nz).
    After the rank transform every row with S >= 2 is flagged non-zero, so
    the flags are set by hand here: 300 of 2,600 rows."""
    w16, unit, anynz, b0 = _inputs(2600, 20, 31)
    anynz = anynz.copy()
    anynz[np.random.default_rng(3).permutation(2600)[:300]] = 0
    _same(w16, unit, anynz, 20, b0, 50, avx)


def test_port_refuses_small_assemblies():
    w16, unit, anynz, b0 = _inputs(2600, 3, 5)
    with pytest.raises(RuntimeError):
        orc.port_fused_graph(w16[:2499], unit[:2499], anynz[:2499], 3, b0)
