"""GPU: the served form.  Every served result must equal the plain build of the
session's whole assembly as of that request (contigs of earlier requests
first), whether the request shared its step with others (the mix) or ran
alone; extensions of fused sessions evaluate only the pairs with a new contig
and merge the session's resident candidate lists."""
import time

import numpy as np
import pytest

from oracle import oracle as orc
from rabbitbin_amd import api, synth

pytestmark = pytest.mark.gpu


def _part(n, genomes, S, seed):
    seqs, off, depth, _ = synth.make_assembly(n, genomes, np.log(8000), 0.8, S, seed=seed)
    return seqs, off, (depth if S else None)


def _concat(parts):
    seqs = np.concatenate([p[0] for p in parts])
    offs, base = [np.zeros(1, np.uint64)], 0
    for p in parts:
        offs.append(p[1][1:] - p[1][0] + np.uint64(base))
        base += int(p[1][-1] - p[1][0])
    depth = None if parts[0][2] is None else np.concatenate([p[2] for p in parts])
    return seqs, np.concatenate(offs), depth


# sessions: (initial (n, genomes, S), extension sizes).  Covers the small
# regime, small -> fused (24,500 + 2,000 > 25,000: a full fused pass), fused ->
# fused (incremental), S = 0 and S = 1, S on both sides of 16 and 32, and a
# fused session without the abundance gate (S = 2: dense counts, extended
# incrementally through its resident count matrix).
SESSIONS = [
    ((3000, 30, 3), [2000]),
    ((26000, 200, 10), [2000, 2000]),
    ((24500, 150, 30), [2000, 2000]),
    ((1000, 10, 0), [500]),
    ((1200, 12, 1), [800]),
    ((27000, 180, 40), [2000]),
    ((25500, 150, 2), [2000]),
]


@pytest.fixture(scope="module")
def workload():
    out = []
    for k, ((n, g, S), ext) in enumerate(SESSIONS):
        parts = [_part(n, g, S, 500 + 10 * k)]
        for e, m in enumerate(ext):
            parts.append(_part(m, max(2, g // 10), S, 501 + 10 * k + e))
        out.append(parts)
    return out


@pytest.fixture(scope="module")
def plain(ctx, workload):
    """Plain builds of every prefix of every session."""
    p = api.default_params(rand_seed=1)
    res = []
    for parts in workload:
        res.append([ctx.build_graph(*_concat(parts[:t + 1]), params=p) for t in range(len(parts))])
    return res


def _same(r, q, what):
    assert r["regime"] == q["regime"], what
    assert r["b0"] == q["b0"], what
    assert r["sim_cutoff"] == q["sim_cutoff"], what
    assert r["gated_pairs"] == q["gated_pairs"], what
    assert r["n_edges"] == q["n_edges"], what
    for k in ("from", "to", "scomp"):
        assert np.array_equal(r[k], q[k]), (what, k)
    assert (r["score"] is None) == (q["score"] is None), what
    if r["score"] is not None:
        assert np.array_equal(r["score"], q["score"]), what


def _serve(workload, max_step_requests):
    """Submits every session at once; each extension is submitted right away
    too, so it queues behind its session's earlier request."""
    results = [[None] * len(parts) for parts in workload]
    infos = [[None] * len(parts) for parts in workload]
    with api.Server(0, max_step_requests=max_step_requests) as srv:
        reqs = []
        for k, parts in enumerate(workload):
            ses, r = srv.submit(*parts[0])
            reqs.append((k, 0, r))
            for t in range(1, len(parts)):
                reqs.append((k, t, srv.extend(ses, *parts[t])))
        for k, t, r in reqs:
            results[k][t], infos[k][t] = srv.wait(r, timeout=300)
        st = srv.stats()
    return results, infos, st


def test_served_mix_equals_plain(workload, plain):
    res, infos, st = _serve(workload, 64)
    assert st["requests"] == sum(len(p) for p in workload)
    assert st["steps"] < st["requests"]  # requests shared steps
    shared = 0
    for k, parts in enumerate(workload):
        for t in range(len(parts)):
            _same(res[k][t], plain[k][t], (k, t))
            inf = infos[k][t]
            assert inf["status"] == 0 and inf["n"] == sum(len(p[1]) - 1 for p in parts[:t + 1])
            shared += inf["step_requests"] > 1
            # an extension is incremental when its session stays in the same regime
            same = t > 0 and plain[k][t - 1]["regime"] == plain[k][t]["regime"]
            assert inf["incremental"] == int(same), (k, t)
    assert shared > 0
    assert any(i["incremental"] for row in infos for i in row)


def test_served_alone_equals_plain(workload, plain):
    res, infos, st = _serve(workload, 1)
    assert st["steps"] == st["requests"]
    for k, parts in enumerate(workload):
        for t in range(len(parts)):
            assert infos[k][t]["step_requests"] == 1
            _same(res[k][t], plain[k][t], (k, t))


def test_served_against_oracle():
    """A small session and its extension against the oracle directly."""
    a = _part(500, 8, 3, 91)
    b = _part(300, 4, 3, 92)
    with api.Server(0) as srv:
        ses, r0 = srv.submit(*a)
        r1 = srv.extend(ses, *b)
        g0, _ = srv.wait(r0, timeout=120)
        g1, _ = srv.wait(r1, timeout=120)
    for g, parts in ((g0, [a]), (g1, [a, b])):
        seqs, off, depth = _concat(parts)
        ref = orc.full_pipeline(seqs, off, depth, rand_seed=1)
        assert g["b0"] == ref["b0"] and g["sim_cutoff"] == ref["sim_cutoff"]
        assert set(zip(g["from"].tolist(), g["to"].tolist())) == set(zip(ref["from"].tolist(),
                                                                        ref["to"].tolist()))
        assert np.array_equal(g["scomp"], ref["sv"])
        assert np.max(np.abs(g["score"].astype(np.float64) - ref["score64"]), initial=0) <= 1e-6


def test_server_budget_and_errors():
    small = _part(1000, 10, 3, 93)
    big = _part(26000, 200, 10, 94)
    with api.Server(0, hbm_budget=400 << 20) as srv:
        sb, rb = srv.submit(*big)
        ses, rs = srv.submit(*small)
        with pytest.raises(api.RbgError, match="status -5"):
            srv.wait(rb, timeout=120)
        g, info = srv.wait(rs, timeout=120)
        assert info["status"] == 0 and g["regime"] == 1
        with pytest.raises(api.RbgError):
            srv.extend(ses, *_part(100, 2, 10, 95))  # sample count differs from the session's
        with pytest.raises(api.RbgError):
            srv.extend(12345, *small)  # unknown session
        srv.release(ses)
        with pytest.raises(api.RbgError):
            srv.extend(ses, *small)  # released
        with pytest.raises(api.RbgError, match="failed earlier"):
            srv.wait(srv.extend(sb, *_part(100, 2, 10, 96)), timeout=60)  # its first request failed
        assert srv.stats()["sessions"] == 1  # the failed session stays until released
        srv.release(sb)
        assert srv.stats()["sessions"] == 0


def test_server_stop_while_busy():
    """Stopping with requests queued, on the device and being copied out
    returns promptly (the worker and the finisher both drain), after a request
    that finished before the stop came out whole."""
    parts = [_part(3000 + 700 * k, 30, 10, 700 + k) for k in range(8)]
    srv = api.Server(0, max_step_requests=2)
    reqs = [srv.submit(*p)[1] for p in parts]
    g, info = srv.wait(reqs[0], timeout=120)
    assert info["status"] == 0 and g["n_edges"] > 0 and info["n"] == 3000
    t0 = time.time()
    srv.stop()
    assert time.time() - t0 < 60

