"""Test helpers for the ingest path: FASTA and depth-file text built from
synthetic assemblies, and a line-based Python model of the reference's FASTA
rules (src/rabbitbin.cpp:1890-2248), written independently of the oracle's
memchr restatement so that each checks the other."""
import numpy as np


def model_fasta(buf, min_contig=2500, min_small=1000):
    """Records (name, class, sequence or None) and num_seqs, by lines: a record
    starts at the first '>' of a line; its name runs to the first byte <= ' ';
    the header ends at the next newline (no newline: the parse stops); the
    body runs to the next record start; bodies shorter than min_small bytes are
    tiny; a body whose bytes after its first newline are all CR/LF keeps its
    first line, otherwise every line is kept; each kept line loses one
    trailing CR."""
    n = len(buf)
    starts = []
    pos = 0
    while pos < n:
        nl = buf.find(b"\n", pos)
        end = n if nl < 0 else nl
        gt = buf.find(b">", pos, end)
        if gt >= 0:
            starts.append(gt)
        pos = end + 1
    out, nseq = [], 0
    for k, s in enumerate(starts):
        e = starts[k + 1] if k + 1 < len(starts) else n
        q = s + 1
        while q < n and buf[q] > 32:
            q += 1
        name = buf[s + 1:q]
        hn = buf.find(b"\n", q)
        if hn < 0:
            break
        body = buf[hn + 1:e]
        nseq += 1
        if len(body) < min_small:
            out.append((name, "tiny", None))
            continue
        first = body.find(b"\n")
        if first < 0 or all(c in b"\r\n" for c in body[first + 1:]):
            line = body if first < 0 else body[:first]
            seq = line[:-1] if line.endswith(b"\r") else line
        else:
            seq = b"".join(x[:-1] if x.endswith(b"\r") else x for x in body.split(b"\n"))
        cls = "large" if len(seq) >= min_contig else ("small" if len(seq) >= min_small else "tiny")
        out.append((name, cls, seq if cls != "tiny" else None))
    return out, nseq


def fasta_text(seqs, offsets, names, seed, widths=(0, 60, 80, 70), crlf=0.1, comment=0.3, blank=0.05):
    """FASTA text of the contigs: each record wrapped at a width drawn from
    `widths` (0 = one line), some with CRLF line ends, header comments and
    blank lines after the sequence."""
    rng = np.random.default_rng(seed)
    parts = []
    n = len(offsets) - 1
    for i in range(n):
        s = bytes(seqs[int(offsets[i]):int(offsets[i + 1])])
        eol = b"\r\n" if rng.random() < crlf else b"\n"
        hdr = b">" + names[i]
        if rng.random() < comment:
            hdr += b" len=%d\tcomment >x" % len(s)
        parts.append(hdr + eol)
        w = int(rng.choice(widths))
        if w == 0:
            parts.append(s + eol)
        else:
            parts.extend(s[k:k + w] + eol for k in range(0, len(s), w))
        if rng.random() < blank:
            parts.append(eol)
    return b"".join(parts)


def depth_text(names, lens, depth, seed, extra_rows=0, dup_rows=0, formats=("%.4f", "%g", "%.6e", "%.2f")):
    """A depth file (contigName, contigLen, totalAvgDepth, then mean / variance
    per sample) for the contigs, rows shuffled, with numbers in several
    formats, `extra_rows` rows of names not in the FASTA and `dup_rows`
    repeated names (a later row with other values)."""
    rng = np.random.default_rng(seed)
    S = depth.shape[1]
    hdr = b"contigName\tcontigLen\ttotalAvgDepth" + b"".join(b"\ts%d.bam\ts%d.bam-var" % (k, k) for k in range(S))
    rows = []
    for i in range(len(names)):
        f = formats[int(rng.integers(len(formats)))].encode()
        vals = b"".join(b"\t" + (f % float(depth[i, k])) + b"\t" + (f % float(rng.random() * 10)) for k in range(S))
        clen = (b"%d" % lens[i]) if rng.random() < 0.8 else (b"%.5e" % lens[i])
        rows.append(names[i] + b"\t" + clen + b"\t" + (f % float(depth[i].sum())) + vals)
    for k in range(extra_rows):
        rows.append(b"extra_%d\t%d\t1.0" % (k, 5000) + b"\t1.5\t0.1" * S)
    order = list(rng.permutation(len(rows)))
    body = [rows[j] for j in order]
    for k in range(dup_rows):  # repeated names after their first row: ignored
        i = int(rng.integers(len(names)))
        body.append(names[i] + b"\t%d\t9.0" % lens[i] + b"\t99.5\t0.0" * S)
    return hdr + b"\n" + b"\n".join(body) + b"\n"


def names_for(n, prefix=b"ctg"):
    return [prefix + b"_%d" % i for i in range(n)]
