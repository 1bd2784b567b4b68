// CPU oracle, part 4: ingest -- the FASTA parse and the depth-file parse.
// TEST INFRASTRUCTURE ONLY.
//
// orc_fasta_parse restates parse_fasta_mmap_parallel (src/rabbitbin.cpp:
// 1890-2248) for one thread over the whole file (the reference's chunks
// start at "\n>" boundaries, where a sequential parse starts a record too, so
// the split does not change the records): memchr to the next '>', the name
// to the first byte <= ' ', the header to the next '\n' (none: stop), the
// body to the next '>' (or the end), the tiny prune on the raw span, the
// single-line test and the multi-line compaction, classification by length.
// Options at their defaults: no tiny collection, names trimmed (they hold no
// blank anyway), sequences kept.
//
// orc_depth_parse restates parse_depth_buf (src/rabbitbin.cpp:2832-2942) for
// one thread, with the C library's strtod / strtof, and the name -> row map
// the reference builds from it (first row of a name wins); orc_depth_match
// fills a row per contig name (src/rabbitbin.cpp:4495-4545).
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

extern "C" {

// Records of the parse, in file order.  cls: 1 tiny, 2 small, 3 large.
// For small and large records: name [name_beg, +name_len) in the file, the
// kept bases appended to `bases` (caller-sized: at most len bytes), their
// offset and length.  Returns the number of records written (<= cap); *nseq =
// the reference's num_seqs.
struct OrcFaRec {
  uint64_t name_beg;
  uint64_t seq_off;  // into bases
  uint64_t seq_len;
  uint32_t name_len;
  uint32_t cls;
};

uint64_t orc_fasta_parse(const char* buf, uint64_t len, uint64_t min_contig, uint64_t min_small, OrcFaRec* rec,
                         uint64_t cap, char* bases, uint64_t* nseq) {
  const char* p = buf;
  const char* end = buf + len;
  uint64_t nr = 0, nb = 0, ns = 0;
  while (p < end) {
    const char* gt = (const char*)std::memchr(p, '>', (size_t)(end - p));
    if (!gt) break;
    p = gt + 1;
    const char* name_start = p;
    while (p < end && (uint8_t)*p > ' ') ++p;
    const char* name_end = p;
    const char* nl = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    if (!nl) break;
    p = nl + 1;
    const char* next_gt = (const char*)std::memchr(p, '>', (size_t)(end - p));
    const char* seq_end = next_gt ? next_gt : end;
    const char* seq_region = p;
    if ((uint64_t)(seq_end - seq_region) < min_small) {  // tiny prune on the raw span
      p = seq_end;
      ++ns;
      if (nr < cap) rec[nr] = OrcFaRec{(uint64_t)(name_start - buf), 0, 0, (uint32_t)(name_end - name_start), 1};
      ++nr;
      continue;
    }
    const char* inner_nl = (const char*)std::memchr(p, '\n', (size_t)(seq_end - p));
    bool single_line;
    uint64_t seq_len;
    if (!inner_nl) {
      single_line = true;
      seq_len = (uint64_t)(seq_end - p);
    } else {
      single_line = true;
      for (const char* q = inner_nl + 1; q < seq_end; ++q)
        if (*q != '\n' && *q != '\r') {
          single_line = false;
          break;
        }
      seq_len = (uint64_t)(inner_nl - p);
    }
    const uint64_t off = nb;
    if (single_line) {
      if (seq_len > 0 && p[seq_len - 1] == '\r') --seq_len;
      std::memcpy(bases + nb, p, seq_len);
      nb += seq_len;
    } else {
      const char* seg = p;
      while (seg < seq_end) {
        const char* seg_nl = (const char*)std::memchr(seg, '\n', (size_t)(seq_end - seg));
        const char* seg_e = seg_nl ? seg_nl : seq_end;
        size_t slen = (size_t)(seg_e - seg);
        if (slen > 0 && seg[slen - 1] == '\r') --slen;
        if (slen > 0) {
          std::memcpy(bases + nb, seg, slen);
          nb += slen;
        }
        if (!seg_nl) break;
        seg = seg_nl + 1;
      }
      seq_len = nb - off;
    }
    p = seq_end;
    ++ns;
    const bool is_large = seq_len >= min_contig;
    const bool is_small = !is_large && seq_len >= min_small;
    if (!is_large && !is_small) nb = off;  // tiny: its copy is dropped
    if (nr < cap)
      rec[nr] = OrcFaRec{(uint64_t)(name_start - buf), (is_large || is_small) ? off : 0,
                         (is_large || is_small) ? seq_len : 0, (uint32_t)(name_end - name_start),
                         is_large ? 3u : (is_small ? 2u : 1u)};
    ++nr;
  }
  *nseq = ns;
  return nr;
}

}  // extern "C"

namespace {

struct DepthEntry {
  std::vector<float> means;
};

}  // namespace

extern "C" {

// The depth file's rows: the name -> means map (first row of a name wins)
// queried for each contig name (names back to back, name i = names[off[i] ..
// off[i+1])); out (n x S) receives the row, found[i] whether there is one.
// min_depth_len: rows with a shorter contigLen are skipped (the reference's
// default prefilter with min_small_contig); 0 keeps all.  Returns the number
// of rows the map holds.
uint64_t orc_depth_parse_match(const char* buf, uint64_t len, uint32_t S, int cv_ext, uint64_t min_depth_len,
                               const char* names, const uint64_t* off, uint64_t n, float* out, uint8_t* found) {
  std::unordered_map<std::string, DepthEntry> map;
  const char* data_end = buf + len;
  const char* nl0 = (const char*)std::memchr(buf, '\n', len);
  const char* p = nl0 ? nl0 + 1 : data_end;
  const char* end = data_end;
  while (p < end) {
    const char* eol = (const char*)std::memchr(p, '\n', (size_t)(end - p));
    const char* line_end = eol ? eol : end;
    const char* q = p;
    p = eol ? eol + 1 : end;
    if (line_end <= q) continue;
    const char* tabp = (const char*)std::memchr(q, '\t', (size_t)(line_end - q));
    if (!tabp) continue;
    std::string name(q, tabp);
    const size_t sp = name.find_first_of(" \t");
    if (sp != std::string::npos) name.resize(sp);
    const char* c = tabp + 1;
    tabp = (const char*)std::memchr(c, '\t', (size_t)(line_end - c));
    if (!tabp) continue;
    char* ep = nullptr;
    const double contig_len_d = std::strtod(c, &ep);
    if (ep != tabp) continue;
    const unsigned long contig_len = (unsigned long)(contig_len_d < 0 ? 0 : contig_len_d);
    c = tabp + 1;
    if (min_depth_len && contig_len < min_depth_len) continue;
    if (!cv_ext) {
      tabp = (const char*)std::memchr(c, '\t', (size_t)(line_end - c));
      if (!tabp) continue;
      c = tabp + 1;
    }
    DepthEntry entry;
    entry.means.assign(S, 0.f);
    auto next_field = [&](const char* cur) -> const char* {
      const char* tp = (const char*)std::memchr(cur, '\t', (size_t)(line_end - cur));
      return tp ? tp + 1 : line_end;
    };
    for (uint32_t i = 0; i < S; ++i) {
      if (c >= line_end) break;
      char* ep2 = nullptr;
      entry.means[i] = std::strtof(c, &ep2);
      c = next_field(c);
      if (!cv_ext) {
        if (c >= line_end) break;
        c = next_field(c);
      }
    }
    map.emplace(std::move(name), std::move(entry));  // keeps the first row of a name
  }
  for (uint64_t i = 0; i < n; ++i) {
    const std::string nm(names + off[i], names + off[i + 1]);
    auto it = map.find(nm);
    found[i] = it != map.end();
    for (uint32_t k = 0; k < S; ++k) out[i * S + k] = found[i] ? it->second.means[k] : 0.f;
  }
  return map.size();
}

// Feature columns of the depth file as ncols(file, 1) counts them on its
// first data line (src/impl/rb_abundance.cpp:50-58: the second line's
// tab-separated fields, a trailing empty field not counted).
uint64_t orc_depth_ncols(const char* buf, uint64_t len) {
  const char* end = buf + len;
  const char* nl0 = (const char*)std::memchr(buf, '\n', len);
  if (!nl0) return 0;
  const char* p = nl0 + 1;
  if (p >= end) return 0;
  const char* nl1 = (const char*)std::memchr(p, '\n', (size_t)(end - p));
  const char* le = nl1 ? nl1 : end;
  uint64_t nc = 0;
  const char* f = p;
  while (f < le) {
    const char* t = (const char*)std::memchr(f, '\t', (size_t)(le - f));
    ++nc;
    if (!t) break;
    f = t + 1;
  }
  return nc;
}

}  // extern "C"
