"""Pins the depth-row arithmetic (unit rows feeding the FP32 abundance gate and
the edge weights, and the column norms of the weighted Jaccard) against a probe
built with the reference's own compiler flags.

The probe (tests/probes/depth_rows_probe.cpp) restates src/rabbitbin.cpp:
4972-4988 and 4799-4806 and src/ranker.h:55-66 in the reference's expression
form.  Built as the reference is built (-std=c++17 -O3 -march=native
-funroll-loops, no contraction flag), its float bits are what the reference
computes on this host; librbg's rbg_depth_features and the oracle must equal
them bit for bit.

Why the contraction of `ss += d * d` cannot matter here: a row of average
ranks always sums to S(S+1)/2 (ties share the mean position), so the mean is
exactly (S+1)/2, every d is a multiple of 1/2 and every d*d and partial sum is
exact in FP64.  Only the square root, the division and the final product
round, all correctly rounded by IEEE 754.  test_rank_rows_are_exact checks the
premise, and a build with -ffp-contract=off must give the same bits.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from rabbitbin_amd import api
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "probes", "depth_rows_probe.cpp")
REF_FLAGS = ["-std=c++17", "-O3", "-DNDEBUG", "-march=native", "-funroll-loops"]


def _build(tmp, name, extra=()):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = os.path.join(tmp, name)
    subprocess.run(["g++", *REF_FLAGS, *extra, "-o", exe, PROBE], check=True)
    return exe


@pytest.fixture(scope="module")
def probes(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("probe"))
    return _build(tmp, "probe_ref"), _build(tmp, "probe_nofma", ["-ffp-contract=off"])


def _run(exe, depth):
    n, S = depth.shape
    inp = np.array([n, S], np.uint64).tobytes() + np.ascontiguousarray(depth, np.float32).tobytes()
    out = subprocess.run([exe], input=inp, stdout=subprocess.PIPE, check=True).stdout
    unit = np.frombuffer(out[:4 * n * S], np.float32).reshape(n, S)
    colnorm = np.frombuffer(out[4 * n * S:], np.float64)
    return unit, colnorm


def _rows(S, seed):
    """Random depths, tie-heavy rows (few distinct values), constant rows,
    all-zero rows and rows with one outlier."""
    rng = np.random.default_rng(seed)
    n = 400
    d = rng.lognormal(0.0, 1.5, (n, S)).astype(np.float32) * 20
    d[50:150] = rng.integers(0, 3, (100, S)).astype(np.float32)        # ties everywhere
    d[150:170] = np.float32(rng.uniform(0, 10))                          # constant rows
    d[170:180] = 0.0
    d[180:200] = rng.integers(0, 2, (20, S)).astype(np.float32) * 1e-3
    d[200:220, 0] = 1e6                                                  # one outlier
    d[220:240] = np.round(d[220:240], 1)
    return d


@pytest.mark.parametrize("S", list(range(2, 65)))
def test_unit_rows_and_colnorm_match_reference_flags(probes, S):
    ref_exe, _ = probes
    d = _rows(S, 1000 + S)
    u_ref, cn_ref = _run(ref_exe, d)
    _, u_lib, _, cn_lib, _ = api.depth_features(d)
    _, u_orc, _, cn_orc, _ = orc.depth_features(d)
    assert np.array_equal(u_lib.view(np.uint32), u_ref.view(np.uint32))
    assert np.array_equal(u_orc.view(np.uint32), u_ref.view(np.uint32))
    assert np.array_equal(cn_lib.view(np.uint64), cn_ref.view(np.uint64))
    assert np.array_equal(cn_orc.view(np.uint64), cn_ref.view(np.uint64))


def test_rank_rows_are_exact(probes):
    """The premise of the argument above, on the same rows: the ranks sum to
    S(S+1)/2 and the squared deviations add up exactly (the FP64 sum equals
    the exact rational sum), and the uncontracted build gives the same bits."""
    from fractions import Fraction
    ref_exe, nofma_exe = probes
    for S in range(2, 65):
        d = _rows(S, 1000 + S)
        ranked, _, _, _, _ = orc.depth_features(d)
        assert np.all(ranked.astype(np.float64).sum(axis=1) == S * (S + 1) / 2)
        for row in ranked[::7]:
            mean = Fraction(S + 1, 2)
            exact = sum((Fraction(float(x)) - mean) ** 2 for x in row)
            fp = 0.0
            for x in row:
                fp += (float(x) - float(mean)) ** 2
            assert Fraction(fp) == exact
        assert np.array_equal(_run(ref_exe, d)[0].view(np.uint32), _run(nofma_exe, d)[0].view(np.uint32))
