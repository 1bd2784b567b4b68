"""RBCACHE1 cache files (rbg_cache_write / rbg_cache_read, host only).

The layout is pinned independently: a writer restated here from
src/impl/rb_cache.cpp:94-175 with struct.pack must produce the library's
bytes exactly, and files it writes in the older versions (1, 2: size_t
endpoints) must read back with the reference's range checks."""
import struct

import numpy as np
import pytest

from rabbitbin_amd import api


def ref_layout(version, f, snv=None):
    """Bytes of an RBCACHE1 file in the layout of src/impl/rb_cache.cpp
    (rbc_wr_pod / rbc_wr_vec / rbc_wr_str / rbc_wr_mat, little endian)."""
    out = [b"RBCACHE1", struct.pack("<III", version, 4, 8)]
    out.append(struct.pack("<QQQQQQddQQB", f["nobs"], f["nobs1"], f["samples"], f["seed"], f["total_size"],
                           f["total_size1"], f["sim_cutoff"], f["b0"], f["min_contig"], f["min_small_contig"],
                           f["has_depth"]))

    def strvec(v):
        b = [struct.pack("<Q", len(v))]
        for s in v:
            e = s.encode()
            b += [struct.pack("<I", len(e)), e]
        return b"".join(b)

    def vec(a, dt):
        a = np.asarray(a, dt)
        return struct.pack("<Q", a.size) + a.astype(np.dtype(dt).newbyteorder("<")).tobytes()

    def mat(m):
        m = np.asarray(m, np.float32)
        return struct.pack("<QQ", *m.shape) + m.astype("<f4").tobytes()

    out += [strvec(f["names"]), strvec(f["small_names"]), vec(f["seq_lens"], np.uint64),
            vec(f["small_seq_lens"], np.uint64), mat(f["depth"]), mat(f["small_depth"]),
            vec(f["large_means"], np.float32), vec(f["depth_raw"], np.float32), vec(f["colnorm"], np.float64)]
    ep = np.uint32 if version >= 3 else np.uint64
    out += [vec(f["from_"], ep), vec(f["to"], ep), vec(f["scomp"], np.float32)]
    if version >= 2:
        if snv is None:
            out.append(b"\0")
        else:
            out += [b"\1", vec(snv[0], np.float32), vec(snv[1], np.uint32), vec(snv[2], np.uint32)]
    return b"".join(out)


def sample(seed=0, nobs=50, nobs1=7, S=4, ne=300):
    rng = np.random.default_rng(seed)
    return {"nobs": nobs, "nobs1": nobs1, "samples": S, "seed": 42 + seed, "total_size": 123456789,
            "total_size1": 4567, "sim_cutoff": 137.25, "b0": 0.0875, "min_contig": 2500, "min_small_contig": 1000,
            "has_depth": 1, "names": [f"g{seed}_c{i}" + "x" * (i % 5) for i in range(nobs)],
            "small_names": [f"s{i}" for i in range(nobs1)],
            "seq_lens": rng.integers(2500, 200000, nobs), "small_seq_lens": rng.integers(1000, 2500, nobs1),
            "depth": rng.random((nobs, S)).astype(np.float32), "small_depth": rng.random((nobs1, S)).astype(np.float32),
            "large_means": rng.random(nobs * S).astype(np.float32), "depth_raw": rng.random(nobs * S).astype(np.float32),
            "colnorm": rng.random(S), "from_": rng.integers(0, nobs, ne), "to": rng.integers(0, nobs, ne),
            "scomp": rng.random(ne).astype(np.float32)}


def test_write_matches_reference_layout(tmp_path):
    f = sample()
    p = str(tmp_path / "g.cache")
    api.write_cache(p, **f)
    assert open(p, "rb").read() == ref_layout(3, f)


def _same(got, f):
    for k in api._CACHE_SCALARS:
        assert got[k] == f[k], k
    assert got["names"] == f["names"] and got["small_names"] == f["small_names"]
    for k in ("seq_lens", "small_seq_lens", "depth", "small_depth", "large_means", "depth_raw", "colnorm", "from_",
              "to", "scomp"):
        assert np.array_equal(got[k], np.asarray(f[k]).astype(got[k].dtype)), k


def test_round_trip_and_older_versions(tmp_path):
    f = sample(1)
    p = str(tmp_path / "g.cache")
    api.write_cache(p, **f)
    got = api.read_cache(p)
    assert got["version"] == 3
    _same(got, f)
    for version in (1, 2):
        q = tmp_path / f"v{version}.cache"
        q.write_bytes(ref_layout(version, f, snv=None))
        got = api.read_cache(str(q))
        assert got["version"] == version
        _same(got, f)
    # a version-2 file with an SNV block
    snv = (np.ones(f["nobs"] * f["samples"], np.float32), np.arange(f["nobs"] * f["samples"]),
           np.arange(f["nobs"] * f["samples"]) * 2)
    q = tmp_path / "snv.cache"
    q.write_bytes(ref_layout(2, f, snv=snv))
    _same(api.read_cache(str(q)), f)


def test_rejects_bad_files(tmp_path):
    f = sample(2)
    good = ref_layout(3, f)
    cases = {"magic": b"RBCACHE2" + good[8:], "version": good[:8] + struct.pack("<I", 9) + good[12:],
             "abi": good[:12] + struct.pack("<I", 8) + good[16:]}
    for name, data in cases.items():
        q = tmp_path / f"{name}.cache"
        q.write_bytes(data)
        with pytest.raises(api.RbgError, match="status -4"):
            api.read_cache(str(q))
    for cut in (20, len(good) // 2, len(good) - 3):  # truncated
        q = tmp_path / f"cut{cut}.cache"
        q.write_bytes(good[:cut])
        with pytest.raises(api.RbgError, match="status -2"):
            api.read_cache(str(q))
    # endpoints beyond nobs, or beyond 32 bits in a version-2 file
    g = dict(f, to=np.full(len(f["to"]), f["nobs"]))
    q = tmp_path / "range.cache"
    q.write_bytes(ref_layout(3, g))
    with pytest.raises(api.RbgError):
        api.read_cache(str(q))
    g = dict(f, from_=np.full(len(f["from_"]), 1 << 33, np.uint64))
    q.write_bytes(ref_layout(2, g))
    with pytest.raises(api.RbgError):
        api.read_cache(str(q))
    with pytest.raises(api.RbgError):  # writer: endpoint out of range
        api.write_cache(str(tmp_path / "w.cache"), **dict(f, to=np.full(len(f["to"]), f["nobs"])))
