"""Canonical digests of graph results, shared by the fixture generator
(tests/golden/make_graph_digests.py, which runs the oracle) and the GPU tests
(which run the library on the same seeded inputs)."""
import hashlib

import numpy as np


def input_digest(seqs, offsets, depth):
    """SHA-256 of an assembly's contig bytes, offsets and depth matrix: the
    GPU tests check it first, so a host whose synthetic input differs is
    reported as such rather than as a graph mismatch."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(seqs, np.uint8).tobytes())
    h.update(np.ascontiguousarray(offsets, np.uint64).astype("<u8").tobytes())
    if depth is not None:
        h.update(np.ascontiguousarray(depth, np.float32).astype("<f4").tobytes())
    return h.hexdigest()


def graph_digest(frm, to, scomp, score):
    """SHA-256 of the edge list sorted by (from, to): endpoints, sComp bits and
    (when the assembly has depth samples) edgeScore bits."""
    frm = np.asarray(frm, np.uint32)
    to = np.asarray(to, np.uint32)
    order = np.lexsort((to, frm))
    h = hashlib.sha256()
    h.update(frm[order].astype("<u4").tobytes())
    h.update(to[order].astype("<u4").tobytes())
    h.update(np.asarray(scomp, np.float32)[order].view(np.uint32).astype("<u4").tobytes())
    if score is not None:
        h.update(np.asarray(score, np.float32)[order].view(np.uint32).astype("<u4").tobytes())
    return h.hexdigest()


def record(frm, to, scomp, score, sim_cutoff, b0, regime, gated_pairs):
    """The comparable summary of one graph."""
    return {"n_edges": int(len(frm)), "sim_cutoff": float(sim_cutoff), "b0": float(b0),
            "regime": int(regime), "gated_pairs": int(gated_pairs),
            "digest": graph_digest(frm, to, scomp, score)}


def record_of_result(g):
    """Summary of a library result dict (Context.build_graph, Batch.result,
    Server.wait)."""
    return record(g["from"], g["to"], g["scomp"], g["score"], g["sim_cutoff"], g["b0"], g["regime"],
                  g["gated_pairs"])


def record_of_oracle(ref, S):
    """Summary of an oracle full_pipeline result; edgeScore is the reference's
    arithmetic (score_ref)."""
    return record(ref["from"], ref["to"], ref["sv"], ref["score_ref"] if S else None, ref["sim_cutoff"],
                  ref["b0"], ref["regime"], ref["gated_pairs"])
