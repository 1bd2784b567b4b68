"""GPU: rbg_edge_weights_noalloc, the edge-weight entry point that allocates
nothing (its staging is a fixed device array), called through the C ABI,
equals rbg_edge_weights bit for bit on two calls in a row, the second with a
smaller input (the array keeps the first call's contents), and refuses
inputs beyond its limit with a clear error."""
import ctypes as C

import numpy as np
import pytest

from rabbitbin_amd import api
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _inputs(rng, n, S, ne):
    d = rng.poisson(rng.lognormal(1, 1, (n, 1)) * rng.lognormal(0, 1, (1, S)) * 4).astype(np.float32)
    _, unit, raw, colnorm, _ = orc.depth_features(d)
    fr = rng.integers(0, n, ne).astype(np.uint32)
    to = rng.integers(0, n, ne).astype(np.uint32)
    sc = rng.random(ne).astype(np.float32)
    return fr, to, sc, unit, raw, colnorm


def _noalloc(ctx, fr, to, sc, unit, raw, colnorm, S):
    L = api.lib()
    L.rbg_edge_weights_noalloc.argtypes = L.rbg_edge_weights.argtypes
    n = unit.shape[0]
    w = np.zeros(len(fr), np.float32)
    w64 = np.zeros(len(fr), np.float64)
    p = api._p
    st = L.rbg_edge_weights_noalloc(ctx._h, p(fr, api.u32p), p(to, api.u32p), p(sc, api.f32p), len(fr),
                                    p(unit, api.f32p), p(raw, api.f32p), p(colnorm, api.f64p), n, S,
                                    orc.w_comp_for(S), 0.7, -0.3, p(w, api.f32p), p(w64, api.f64p))
    return st, w, w64


def test_noalloc_equals_edge_weights(ctx):
    rng = np.random.default_rng(11)
    for n, S, ne in ((20000, 30, 600000), (3000, 16, 40000)):  # second call smaller
        fr, to, sc, unit, raw, colnorm = _inputs(rng, n, S, ne)
        st, w, w64 = _noalloc(ctx, fr, to, sc, unit, raw, colnorm, S)
        assert st == 0, api.lib().rbg_last_error(ctx._h)
        rw, rw64 = ctx.edge_weights(fr, to, sc, unit, raw, colnorm, S, orc.w_comp_for(S))
        assert np.array_equal(w.view(np.uint32), rw.view(np.uint32))
        assert np.array_equal(w64.view(np.uint64), rw64.view(np.uint64))


def test_noalloc_refuses_beyond_its_limit(ctx):
    rng = np.random.default_rng(12)
    fr, to, sc, unit, raw, colnorm = _inputs(rng, 1000, 3, 3_000_000)  # 72 MB of edge staging
    st, _, _ = _noalloc(ctx, fr, to, sc, unit, raw, colnorm, 3)
    assert st == -4
    assert b"limit is 64 MiB" in api.lib().rbg_last_error(ctx._h)
