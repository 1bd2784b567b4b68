"""Ingest (FASTA and depth-file parsing) on the CPU: the oracle's restatement
of the reference's parsers pinned by the reference's own test data
(test/contigs.fa, test/contigs_depth.txt, copied to tests/golden/), by
hand-worked cases, and by an independent line-based model of the FASTA rules
on random byte soups; and librbg's decimal conversion (the depth parser's
strtof / strtod) against the C library, compiled for the host."""
import os
import subprocess

import numpy as np
import pytest

from ingest_text import depth_text, fasta_text, model_fasta, names_for
from oracle import oracle as orc
from rabbitbin_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def test_reference_fixture():
    fa = open(os.path.join(GOLD, "ref_contigs.fa"), "rb").read()
    recs, nseq = orc.fasta_parse(fa)
    assert nseq == 1 and len(recs) == 1
    name, cls, seq = recs[0]
    # the header's comment is cut at the first blank; the depth file of the
    # reference's own test gives the same contig 5,404 bases
    assert name == b"NODE_1_length_5374_cov_8.558988" and cls == "large" and len(seq) == 5404
    assert seq == b"".join(fa.split(b"\n")[1:]) and set(seq) <= set(b"ACGT")
    assert model_fasta(fa) == (recs, nseq)
    dp = open(os.path.join(GOLD, "ref_contigs_depth.txt"), "rb").read()
    assert orc.depth_ncols(dp) == 5  # name, contigLen, totalAvgDepth, mean, variance: S = 1
    m, found, rows = orc.depth_parse_match(dp, [name], 1)
    assert found[0] and rows == 1
    assert m[0, 0].view(np.uint32) == 0x416373EB  # strtof("14.2158")


HAND = [
    # (text, expected records as (name, class, seq)); min_contig 4, min_small 2
    (b">a\nACGT\n", [(b"a", "large", b"ACGT")]),
    (b">a desc\nAC\nGT\n>b\nA\n", [(b"a", "large", b"ACGT"), (b"b", "tiny", None)]),
    (b">a\r\nACGT\r\n\r\n", [(b"a", "large", b"ACGT")]),                      # single line, CR dropped
    (b">a\nAC\r\nGT\r\n", [(b"a", "large", b"ACGT")]),                        # multi-line, CRs dropped
    (b">a\nAC\r\r\nGT\n", [(b"a", "large", b"AC\rGT")]),                     # one CR per line end only
    (b">a\nACGT\n\r\r\n", [(b"a", "large", b"ACGT")]),                        # trailing CR/LF lines ignored
    (b">a\nAC\n\nGT\n", [(b"a", "large", b"ACGT")]),                          # blank line inside: multi-line
    (b">a>b\nACGT\n", [(b"a>b", "large", b"ACGT")]),                          # '>' inside a header
    (b">a\nACG>b\nTTTT\n", [(b"a", "small", b"ACG"), (b"b", "large", b"TTTT")]),  # '>' mid-line starts a record
    (b"junk\n>a\nACGT", [(b"a", "large", b"ACGT")]),                          # text before the first '>'; no final LF
    (b">a\nACGT\n>b", [(b"a", "large", b"ACGT")]),                            # header without a newline: stop
    (b">\nACGT\n", [(b"", "large", b"ACGT")]),                                # empty name
    (b">a\tx\nAC\n", [(b"a", "small", b"AC")]),
    (b">a\nA\n", [(b"a", "tiny", None)]),                                     # raw body span 2 >= 2, seq 1
    (b"", []),
    (b"no record\n", []),
]


@pytest.mark.parametrize("text,want", HAND)
def test_fasta_hand_cases(text, want):
    recs, nseq = orc.fasta_parse(text, min_contig=4, min_small=2)
    assert recs == want
    assert nseq == len(want)
    assert model_fasta(text, 4, 2) == (recs, nseq)


def test_fasta_random_soups_model_equals_oracle():
    rng = np.random.default_rng(17)
    alphabet = np.frombuffer(b"ACGTACGTACGTACGTN\n\n\n>\r \ta", np.uint8)
    for case in range(3000):
        n = int(rng.integers(0, 300))
        buf = alphabet[rng.integers(0, len(alphabet), n)].tobytes()
        mc, ms = int(rng.integers(1, 12)), int(rng.integers(0, 6))
        got = orc.fasta_parse(buf, mc, ms)
        assert got == model_fasta(buf, mc, ms), (case, buf)


def test_fasta_of_a_synthetic_assembly():
    seqs, off, _, _ = synth.make_assembly(300, 10, np.log(3000), 1.0, 0, seed=5)
    names = names_for(300)
    text = fasta_text(seqs, off, names, seed=1)
    recs, nseq = orc.fasta_parse(text, min_contig=2500, min_small=1000)
    assert (recs, nseq) == model_fasta(text, 2500, 1000)
    assert nseq == 300
    for i, (name, cls, s) in enumerate(recs):
        want = bytes(seqs[int(off[i]):int(off[i + 1])])
        assert name == names[i]
        L = len(want)
        assert cls == ("large" if L >= 2500 else "small" if L >= 1000 else "tiny")
        if cls != "tiny":
            assert s == want


def test_depth_hand_cases():
    text = (b"contigName\tcontigLen\ttotalAvgDepth\ts1\ts1-var\ts2\ts2-var\n"
            b"a\t5000\t3.0\t1.5\t0.1\t2.25\t0.2\n"
            b"b x\t2.5e3\t1\t7\t1\t8\t1\n"          # name trimmed at the blank; contigLen in %e form
            b"c\t12abc\t1\t7\t1\t8\t1\n"            # contigLen not a whole field: row skipped
            b"d\t500\t1\t7\t1\t8\t1\n"              # below the prefilter: skipped
            b"e\t5000\t1\t4.5\t1\n"                 # short row: the missing mean stays 0
            b"a\t5000\t3.0\t9\t9\t9\t9\n"           # repeated name: the first row wins
            b"f\t5000\t1\t\t1\t6\t1\n"              # empty mean field: strtof skips the tab and reads the next field
            b"\n"
            b"g\t5000\n"                            # no totalAvgDepth: skipped
            b"h\t5000\t1\t-0.5e1\t1\tinf\t1\r\n")   # sign, exponent, inf; CR before the newline
    names = [b"a", b"b", b"c", b"d", b"e", b"f", b"g", b"h"]
    m, found, rows = orc.depth_parse_match(text, names, 2, min_depth_len=1000)
    assert list(found) == [True, True, False, False, True, True, False, True]
    assert rows == 5
    np.testing.assert_array_equal(m[0], np.float32([1.5, 2.25]))
    np.testing.assert_array_equal(m[1], np.float32([7, 8]))
    np.testing.assert_array_equal(m[4], np.float32([4.5, 0]))
    np.testing.assert_array_equal(m[5], np.float32([1, 6]))
    np.testing.assert_array_equal(m[7], np.float32([-5, np.inf]))
    assert orc.depth_ncols(text) == 7


def test_depth_of_a_synthetic_assembly():
    rng = np.random.default_rng(3)
    n, S = 200, 5
    names = names_for(n)
    lens = rng.integers(500, 20000, n)
    depth = (rng.lognormal(2.0, 1.5, (n, S))).astype(np.float32)
    text = depth_text(names, lens, depth, seed=2, extra_rows=20, dup_rows=10)
    m, found, rows = orc.depth_parse_match(text, names, S, min_depth_len=1000)
    keep = lens >= 1000
    assert np.array_equal(found, keep)
    assert rows == keep.sum() + 20
    # the values written at 4-6 significant digits read back within their rounding
    np.testing.assert_allclose(m[keep], depth[keep], rtol=5e-3, atol=5e-3)


def test_decimal_conversion_matches_the_c_library(tmp_path):
    exe = tmp_path / "decimal_probe"
    src = os.path.join(HERE, "probes", "decimal_probe.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), src], check=True)
    out = subprocess.run([str(exe), "200000"], capture_output=True, text=True, check=True).stdout.strip()
    assert out.startswith("ok "), out
    checked, refused = map(int, out.split()[1:])
    assert checked > 200000 and refused < 0.12 * 2 * checked
