// librbg: ingest on the device side of the boundary -- the FASTA records and
// the depth file's rows, parsed in HBM (gfx950).
//
// FASTA (parse_fasta_mmap_parallel, src/rabbitbin.cpp:1890-2248, default
// options).  A record starts at every '>' that is the first '>' of its line
// (the parser's memchr for '>' from the end of the previous record's body
// finds exactly those: a '>' later on a header line lies inside that header,
// and the body ends at the next '>').  Its name runs from the byte after the
// '>' to the first byte <= ' '; the header ends at the first '\n' after the
// name (a record without one is dropped, and nothing after it can start a
// record); the body runs from there to the next record start.  Bodies shorter
// than min_small_contig bytes (newlines included) are tiny.  The sequence:
// when every byte after the body's first '\n' is '\n' or '\r' (single-line),
// the first line less one trailing '\r'; otherwise every line, each less one
// trailing '\r', concatenated.  Records of at least min_contig bases are
// large, of at least min_small_contig small.
//
// Byte scans use 4096-byte blocks (256 threads x 16 bytes).  Record starts
// need the kind of the last '\n' / '>' before each '>': per block the kind of
// its last such byte and the number of starts it holds when that kind is not
// '>' (one fewer when its first such byte is a '>' and the kind is '>'); a
// look-back gives each block its carry, a scan the positions.
//
// Depth rows (parse_depth_buf, src/rabbitbin.cpp:2832-2942): one thread per
// line, strtod / strtof emulated exactly by rbg_decimal.h.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rbg_decimal.h"
#include "rbg_ingest.h"

namespace rbg {
namespace {

constexpr uint32_t kIgThreads = 256;
constexpr uint32_t kIgBytes = 16;                        // per thread
constexpr uint64_t kIgBlock = kIgThreads * kIgBytes;     // 4096 bytes per block

// kinds of the last line/record byte: 0 none, 1 '\n', 2 '>'
__device__ __forceinline__ uint32_t combine_kind(uint32_t left, uint32_t right) { return right ? right : left; }

// Block-wide exclusive scan of `kind` (last non-zero of the threads before).
__device__ uint32_t block_kind_scan(uint32_t kind, uint32_t* sh) {
  const uint32_t tid = threadIdx.x;
  sh[tid] = kind;
  __syncthreads();
  for (uint32_t d = 1; d < kIgThreads; d <<= 1) {
    const uint32_t v = tid >= d ? sh[tid - d] : 0u;
    __syncthreads();
    sh[tid] = combine_kind(v, sh[tid]);
    __syncthreads();
  }
  const uint32_t excl = tid ? sh[tid - 1] : 0u;
  __syncthreads();
  return excl;
}

// Block-wide exclusive sum of v; *total = the block's sum.
__device__ uint32_t block_sum_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  const uint32_t tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (uint32_t d = 1; d < kIgThreads; d <<= 1) {
    const uint32_t u = tid >= d ? sh[tid - d] : 0u;
    __syncthreads();
    sh[tid] += u;
    __syncthreads();
  }
  const uint32_t incl = sh[tid];
  *total = sh[kIgThreads - 1];
  __syncthreads();
  return incl - v;
}

// The thread's 16 bytes [pos, pos + 16) clipped to len (missing bytes are 0).
__device__ __forceinline__ void load16(const uint8_t* __restrict__ d, uint64_t pos, uint64_t len, uint8_t (&b)[16]) {
  if (pos + 16 <= len) {
    const uint4 v = *reinterpret_cast<const uint4*>(d + pos);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 16; ++k) b[k] = (uint8_t)(w[k >> 2] >> (8 * (k & 3)));
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) b[k] = pos + k < len ? d[pos + k] : (uint8_t)0;
  }
}

// Per thread: kind of its first and last '\n' / '>' and its starts given that
// the kind before it is not '>'.
__device__ __forceinline__ void thread_starts(const uint8_t (&b)[16], uint32_t& first, uint32_t& last,
                                              uint32_t& starts, uint32_t carry_in, uint32_t* where) {
  first = 0;
  last = 0;
  starts = 0;
  uint32_t prev = carry_in;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t kind = b[k] == '\n' ? 1u : (b[k] == '>' ? 2u : 0u);
    if (kind == 2u && prev != 2u) {
      if (where) where[starts] = (uint32_t)k;
      ++starts;
    }
    if (kind) {
      if (!first) first = kind;
      last = kind;
      prev = kind;
    }
  }
}

__global__ __launch_bounds__(kIgThreads) void fa_summary_kernel(const uint8_t* __restrict__ d, uint64_t len,
                                                                 uint32_t* __restrict__ summ) {
  __shared__ uint32_t sh[kIgThreads];
  __shared__ uint32_t first_sh[kIgThreads];
  const uint64_t blk = blockIdx.x;
  const uint64_t pos = blk * kIgBlock + (uint64_t)threadIdx.x * kIgBytes;
  uint8_t b[16];
  load16(d, pos, len, b);
  uint32_t first, last, st;
  thread_starts(b, first, last, st, 0u, nullptr);
  // first kind of the block: of the first thread with any
  first_sh[threadIdx.x] = first;
  const uint32_t before = block_kind_scan(last, sh);  // kind before this thread (0: before the block)
  // starts with the block carry not '>': a thread whose carry is '>' (from a
  // thread before it) loses its first '>' if that is its first special byte
  const uint32_t cnt = st - ((before == 2u && first == 2u) ? 1u : 0u);
  uint32_t total;
  (void)block_sum_scan(cnt, sh, &total);
  if (threadIdx.x == 0) {
    uint32_t f = 0;
    for (uint32_t t = 0; t < kIgThreads && !f; ++t) f = first_sh[t];
    summ[3 * blk + 1] = f;
    summ[3 * blk + 2] = total;
  }
  if (threadIdx.x == kIgThreads - 1) summ[3 * blk + 0] = combine_kind(before, last);
}

// Per group of kIgGroup blocks: the kind of its last '\n' / '>' (0: none).
constexpr uint64_t kIgGroup = 1024;
__global__ void fa_group_kernel(const uint32_t* __restrict__ summ, uint64_t nblk, uint32_t* __restrict__ glast) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t b0 = g * kIgGroup;
  if (b0 >= nblk) return;
  const uint64_t b1 = b0 + kIgGroup < nblk ? b0 + kIgGroup : nblk;
  uint32_t k = 0;
  for (uint64_t b = b0; b < b1; ++b) k = combine_kind(k, summ[3 * b]);
  glast[g] = k;
}

// Per block: the kind before it (look-back within its group, then over the
// groups before: at most kIgGroup + nblk / kIgGroup steps) and its number of
// record starts.
__global__ void fa_carry_kernel(const uint32_t* __restrict__ summ, const uint32_t* __restrict__ glast, uint64_t nblk,
                                uint32_t* __restrict__ carry, uint32_t* __restrict__ count) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  uint32_t c = 0;
  const uint64_t g = b / kIgGroup;
  for (uint64_t q = b; q-- > g * kIgGroup;) {
    const uint32_t k = summ[3 * q];
    if (k) {
      c = k;
      break;
    }
  }
  if (!c)
    for (uint64_t h = g; h-- > 0;)
      if (glast[h]) {
        c = glast[h];
        break;
      }
  carry[b] = c;
  count[b] = summ[3 * b + 2] - ((c == 2u && summ[3 * b + 1] == 2u) ? 1u : 0u);
}

__global__ __launch_bounds__(kIgThreads) void fa_starts_kernel(const uint8_t* __restrict__ d, uint64_t len,
                                                               const uint32_t* __restrict__ carry,
                                                               const uint64_t* __restrict__ off, uint64_t R,
                                                               uint64_t* __restrict__ starts) {
  __shared__ uint32_t sh[kIgThreads];
  const uint64_t blk = blockIdx.x;
  const uint64_t pos = blk * kIgBlock + (uint64_t)threadIdx.x * kIgBytes;
  uint8_t b[16];
  load16(d, pos, len, b);
  uint32_t first, last, st;
  thread_starts(b, first, last, st, 0u, nullptr);
  const uint32_t before = combine_kind(carry[blk], block_kind_scan(last, sh));
  uint32_t where[16];
  thread_starts(b, first, last, st, before, where);
  uint32_t total;
  const uint32_t at = block_sum_scan(st, sh, &total);
  for (uint32_t k = 0; k < st; ++k) {
    const uint64_t i = off[blk] + at + k;
    if (i < R) starts[i] = pos + where[k];  // always: the counts are the summary's
  }
}

// ---------------------------------------------------------------------------
// Records: one wave per record, 256-byte windows (4 bytes per lane).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

// First position in [from, to) whose byte satisfies pred (to when none).
template <class Pred>
__device__ uint64_t wave_find(const uint8_t* __restrict__ d, uint64_t from, uint64_t to, Pred pred) {
  for (uint64_t w = from; w < to; w += 256) {
    const uint64_t p = w + 4 * (uint64_t)lane();
    uint32_t hit = 4;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (hit == 4 && p + k < to && pred(d[p + k])) hit = k;
    const uint64_t m = __ballot(hit < 4);
    if (m) {
      const uint32_t src = (uint32_t)__ffsll((unsigned long long)m) - 1u;
      const uint32_t h = __shfl(hit, (int)src);
      return w + 4 * (uint64_t)src + h;
    }
  }
  return to;
}

__global__ __launch_bounds__(256) void fa_records_kernel(const uint8_t* __restrict__ d, uint64_t len,
                                                         const uint64_t* __restrict__ starts, uint32_t R,
                                                         uint64_t min_contig, uint64_t min_small,
                                                         FaRecord* __restrict__ rec) {
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  const uint64_t s = starts[r];
  const uint64_t e = r + 1 < R ? starts[r + 1] : len;
  FaRecord out{};
  out.name_beg = s + 1;
  const uint64_t name_end = wave_find(d, s + 1, len, [](uint8_t c) { return c <= (uint8_t)' '; });
  out.name_len = (uint32_t)(name_end - (s + 1));
  const uint64_t nl = wave_find(d, name_end, len, [](uint8_t c) { return c == (uint8_t)'\n'; });
  if (nl >= len || nl >= e) {  // header without a newline: dropped (only the last record)
    out.cls = kFaDropped;
    if (lane() == 0) rec[r] = out;
    return;
  }
  const uint64_t p = nl + 1;
  out.body_beg = p;
  out.body_end = e;
  if (e - p < min_small) {
    out.cls = kFaTiny;
    if (lane() == 0) rec[r] = out;
    return;
  }
  // one pass over the body: the first '\n', whether anything but '\n' / '\r'
  // follows it, and the bases kept by the multi-line rule
  uint64_t inner = e;
  bool other_after = false;
  uint64_t kept = 0;
  for (uint64_t w = p; w < e; w += 256) {
    const uint64_t q = w + 4 * (uint64_t)lane();
    uint32_t keep = 0, nlpos = 4, oth = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint64_t x = q + k;
      if (x < e) {
        const uint8_t c = d[x];
        const uint8_t nx = x + 1 < e ? d[x + 1] : (uint8_t)'\n';  // the body's end acts as a line end
        if (c == '\n') {
          if (nlpos == 4) nlpos = k;
        } else if (!(c == '\r' && nx == '\n')) {
          ++keep;
        }
        if (c != '\n' && c != '\r') oth |= 1u << k;
      }
    }
    // first '\n' of the window, if the body's first is not found yet
    const uint64_t m = __ballot(nlpos < 4);
    uint64_t inner_w = e;
    if (m) {
      const uint32_t src = (uint32_t)__ffsll((unsigned long long)m) - 1u;
      inner_w = w + 4 * (uint64_t)src + (uint64_t)__shfl((int)nlpos, (int)src);
    }
    // bytes other than '\n' / '\r' after the body's first '\n'
    const uint64_t first_nl = inner < e ? inner : inner_w;
    bool o = false;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (((oth >> k) & 1u) && q + k > first_nl && first_nl < e) o = true;
    if (__ballot(o)) other_after = true;
    if (inner == e) inner = inner_w;
    // wave sum of keep
    uint32_t sum = keep;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    kept += sum;
  }
  uint64_t len_seq;
  bool single = !other_after;
  if (single) {
    const uint64_t line_end = inner;  // e when the body has no '\n'
    len_seq = line_end - p;
    if (len_seq > 0 && d[line_end - 1] == '\r') --len_seq;
  } else {
    len_seq = kept;
  }
  out.single = single ? 1 : 0;
  out.seq_len = len_seq;
  out.cls = len_seq >= min_contig ? kFaLarge : (len_seq >= min_small ? kFaSmall : kFaTiny);
  if (lane() == 0) rec[r] = out;
}

// Copy the kept bases of the records of class `cls` to out + off[slot[r]].
__global__ __launch_bounds__(256) void fa_copy_kernel(const uint8_t* __restrict__ d, const FaRecord* __restrict__ rec,
                                                      uint32_t R, uint32_t cls, const uint32_t* __restrict__ slot,
                                                      const uint64_t* __restrict__ off, uint8_t* __restrict__ out) {
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= R) return;
  const FaRecord x = rec[r];
  if (x.cls != cls) return;
  const uint64_t base = off[slot[r]];
  const uint64_t p = x.body_beg, e = x.single ? x.body_beg + x.seq_len : x.body_end;
  const uint64_t below = (1ull << lane()) - 1ull;
  uint64_t at = 0;
  for (uint64_t w = p; w < e; w += 256) {
    const uint64_t q = w + 4 * (uint64_t)lane();
    uint8_t c[4];
    bool keep[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint64_t xx = q + k;
      c[k] = xx < e ? d[xx] : (uint8_t)0;
      if (x.single) {
        keep[k] = xx < e;
      } else {
        const uint8_t nx = xx + 1 < e ? d[xx + 1] : (uint8_t)'\n';
        keep[k] = xx < e && c[k] != '\n' && !(c[k] == '\r' && nx == '\n');
      }
    }
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint64_t bk = __ballot(keep[k]);
      before += (uint32_t)__popcll(bk & below);
      total += (uint32_t)__popcll(bk);
    }
    uint32_t own = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k)
      if (keep[k]) out[base + at + before + own++] = c[k];
    at += total;
  }
}

// Gather the names of class-`cls` records into names + name_off[slot[r]].
__global__ void fa_names_kernel(const uint8_t* __restrict__ d, const FaRecord* __restrict__ rec, uint32_t R,
                                uint32_t cls, const uint32_t* __restrict__ slot, const uint64_t* __restrict__ noff,
                                uint8_t* __restrict__ names) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const FaRecord x = rec[r];
  if (x.cls != cls) return;
  const uint64_t o = noff[slot[r]];
  for (uint32_t k = 0; k < x.name_len; ++k) names[o + k] = d[x.name_beg + k];
}

// ---------------------------------------------------------------------------
// Depth file: newline positions, then one thread per line.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kIgThreads) void nl_count_kernel(const uint8_t* __restrict__ d, uint64_t len,
                                                              uint32_t* __restrict__ count) {
  __shared__ uint32_t sh[kIgThreads];
  const uint64_t pos = (uint64_t)blockIdx.x * kIgBlock + (uint64_t)threadIdx.x * kIgBytes;
  uint8_t b[16];
  load16(d, pos, len, b);
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) n += (pos + k < len && b[k] == '\n') ? 1u : 0u;
  uint32_t total;
  (void)block_sum_scan(n, sh, &total);
  if (threadIdx.x == 0) count[blockIdx.x] = total;
}

__global__ __launch_bounds__(kIgThreads) void nl_write_kernel(const uint8_t* __restrict__ d, uint64_t len,
                                                              const uint64_t* __restrict__ off,
                                                              uint64_t* __restrict__ nlpos) {
  __shared__ uint32_t sh[kIgThreads];
  const uint64_t pos = (uint64_t)blockIdx.x * kIgBlock + (uint64_t)threadIdx.x * kIgBytes;
  uint8_t b[16];
  load16(d, pos, len, b);
  uint32_t n = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) n += (pos + k < len && b[k] == '\n') ? 1u : 0u;
  uint32_t total;
  uint64_t at = off[blockIdx.x] + block_sum_scan(n, sh, &total);
  for (int k = 0; k < 16; ++k)
    if (pos + k < len && b[k] == '\n') nlpos[at++] = pos + k;
}

__device__ __forceinline__ uint64_t find_byte(const uint8_t* __restrict__ d, uint64_t from, uint64_t to, uint8_t c) {
  for (uint64_t x = from; x < to; ++x)
    if (d[x] == c) return x;
  return to;
}

// Line l of the data (after the header): [beg, end) with beg = 1 + the l-th
// newline, end = the next newline or the file's end.
__global__ void depth_rows_kernel(const uint8_t* __restrict__ d, uint64_t len, const uint64_t* __restrict__ nlpos,
                                  uint64_t n_nl, uint64_t n_lines, uint32_t S, uint32_t cv_ext,
                                  uint64_t min_depth_len, DepthRow* __restrict__ rows, float* __restrict__ means,
                                  uint32_t* __restrict__ flags) {
  const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lines) return;
  const uint64_t q = nlpos[l] + 1;
  const uint64_t line_end = l + 1 < n_nl ? nlpos[l + 1] : len;
  DepthRow row{};
  row.kept = 0;
  float* mrow = means + l * (uint64_t)S;
  for (uint32_t i = 0; i < S; ++i) mrow[i] = 0.f;
  if (line_end > q) {
    const uint64_t tab = find_byte(d, q, line_end, '\t');
    if (tab < line_end) {
      // name, trimmed at its first ' ' (the first '\t' ends it anyway)
      const uint64_t sp = find_byte(d, q, tab, ' ');
      row.name_beg = q;
      row.name_len = (uint32_t)(sp - q);
      uint64_t c = tab + 1;
      const uint64_t tab2 = find_byte(d, c, line_end, '\t');
      if (tab2 < line_end) {
        uint64_t bits, used;
        const int st = dec::parse_decimal(d + c, d + len, 53, bits, &used);
        if (st != dec::kOk) {
          atomicOr(flags, kDepthUnsupportedNumber);
        } else if (c + used == tab2) {
          const double v = __longlong_as_double((long long)bits);
          bool ok = true;
          uint64_t clen = 0;
          if (v != v || v >= 18446744073709551616.0) {
            atomicOr(flags, kDepthUnsupportedNumber);  // the reference's cast is undefined there
            ok = false;
          } else {
            clen = v < 0 ? 0ull : (uint64_t)v;
          }
          row.contig_len = clen;
          c = tab2 + 1;
          if (ok && !(min_depth_len && clen < min_depth_len)) {
            bool have = true;
            if (!cv_ext) {  // totalAvgDepth
              const uint64_t t3 = find_byte(d, c, line_end, '\t');
              if (t3 >= line_end) have = false;
              c = t3 + 1;
            }
            if (have) {
              row.kept = 1;
              for (uint32_t i = 0; i < S; ++i) {
                if (c >= line_end) break;
                uint64_t fb, fu;
                if (dec::parse_decimal(d + c, d + len, 24, fb, &fu) != dec::kOk) {
                  atomicOr(flags, kDepthUnsupportedNumber);
                  break;
                }
                mrow[i] = __uint_as_float((uint32_t)fb);
                uint64_t t = find_byte(d, c, line_end, '\t');
                c = t < line_end ? t + 1 : line_end;
                if (!cv_ext) {
                  if (c >= line_end) break;
                  t = find_byte(d, c, line_end, '\t');
                  c = t < line_end ? t + 1 : line_end;
                }
              }
            }
          }
        }
      }
    }
  }
  rows[l] = row;
}

// FNV-1a 64 of a name; 0 is kept for empty slots.
__device__ __forceinline__ uint64_t name_hash(const uint8_t* __restrict__ p, uint32_t n) {
  uint64_t h = 1469598103934665603ull;
  for (uint32_t k = 0; k < n; ++k) h = (h ^ p[k]) * 1099511628211ull;
  return h ? h : 1ull;
}

// Insert the kept rows: slot key = hash, value = the smallest line (the first
// row of a name wins, as the reference's map keeps its first insertion).
__global__ void depth_insert_kernel(const uint8_t* __restrict__ d, const DepthRow* __restrict__ rows, uint64_t n_lines,
                                    unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals, uint64_t mask) {
  const uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lines || !rows[l].kept) return;
  const uint64_t h = name_hash(d + rows[l].name_beg, rows[l].name_len);
  for (uint64_t s = h & mask;; s = (s + 1) & mask) {
    const unsigned long long prev = atomicCAS(&keys[s], 0ull, (unsigned long long)h);
    if (prev == 0ull || prev == h) {
      atomicMin(&vals[s], (uint32_t)l);
      return;
    }
  }
}

// Depth row of each contig (names in names/noff): copies its S means into
// out + i*S; a contig without a row sets kDepthMissing and records its index.
__global__ void depth_lookup_kernel(const uint8_t* __restrict__ d, const DepthRow* __restrict__ rows,
                                    const float* __restrict__ means, uint32_t S,
                                    const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ vals,
                                    uint64_t mask, const uint8_t* __restrict__ names, const uint64_t* __restrict__ noff,
                                    uint32_t n, float* __restrict__ out, uint32_t* __restrict__ flags,
                                    uint32_t* __restrict__ first_missing) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t* nm = names + noff[i];
  const uint32_t nlen = (uint32_t)(noff[i + 1] - noff[i]);
  const uint64_t h = name_hash(nm, nlen);
  uint32_t line = 0xFFFFFFFFu;
  for (uint64_t s = h & mask;; s = (s + 1) & mask) {
    const unsigned long long k = keys[s];
    if (k == 0ull) break;
    if (k == h) {
      line = vals[s];
      break;
    }
  }
  bool ok = line != 0xFFFFFFFFu;
  if (ok) {
    const DepthRow r = rows[line];
    if (r.name_len != nlen) {
      ok = false;
    } else {
      for (uint32_t k = 0; k < nlen; ++k)
        if (d[r.name_beg + k] != nm[k]) ok = false;
    }
    if (!ok) atomicOr(flags, kDepthHashCollision);
  }
  float* o = out + (uint64_t)i * S;
  if (!ok) {
    atomicOr(flags, kDepthMissing);
    atomicMin(first_missing, i);
    for (uint32_t k = 0; k < S; ++k) o[k] = 0.f;
    return;
  }
  const float* m = means + (uint64_t)line * S;
  for (uint32_t k = 0; k < S; ++k) o[k] = m[k];
}

__global__ void fill_u32_ig_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
uint64_t ig_blocks(uint64_t len) { return (len + kIgBlock - 1) / kIgBlock; }

hipError_t launch_fa_summary(const uint8_t* d, uint64_t len, uint32_t* summ, hipStream_t st) {
  const uint64_t nb = ig_blocks(len);
  if (!nb) return hipSuccess;
  hipLaunchKernelGGL(fa_summary_kernel, dim3((uint32_t)nb), dim3(kIgThreads), 0, st, d, len, summ);
  return hipGetLastError();
}

uint64_t ig_groups(uint64_t nblk) { return (nblk + kIgGroup - 1) / kIgGroup; }

hipError_t launch_fa_carry(const uint32_t* summ, uint64_t nblk, uint32_t* glast, uint32_t* carry, uint32_t* count,
                           hipStream_t st) {
  if (!nblk) return hipSuccess;
  const uint64_t ng = ig_groups(nblk);
  hipLaunchKernelGGL(fa_group_kernel, dim3((uint32_t)((ng + 255) / 256)), dim3(256), 0, st, summ, nblk, glast);
  hipLaunchKernelGGL(fa_carry_kernel, dim3((uint32_t)((nblk + 255) / 256)), dim3(256), 0, st, summ, glast, nblk,
                     carry, count);
  return hipGetLastError();
}

hipError_t launch_fa_starts(const uint8_t* d, uint64_t len, const uint32_t* carry, const uint64_t* off, uint64_t R,
                            uint64_t* starts, hipStream_t st) {
  const uint64_t nb = ig_blocks(len);
  if (!nb) return hipSuccess;
  hipLaunchKernelGGL(fa_starts_kernel, dim3((uint32_t)nb), dim3(kIgThreads), 0, st, d, len, carry, off, R, starts);
  return hipGetLastError();
}

hipError_t launch_fa_records(const uint8_t* d, uint64_t len, const uint64_t* starts, uint32_t R, uint64_t min_contig,
                             uint64_t min_small, FaRecord* rec, hipStream_t st) {
  if (!R) return hipSuccess;
  hipLaunchKernelGGL(fa_records_kernel, dim3((R + 3) / 4), dim3(256), 0, st, d, len, starts, R, min_contig, min_small,
                     rec);
  return hipGetLastError();
}

hipError_t launch_fa_copy(const uint8_t* d, const FaRecord* rec, uint32_t R, uint32_t cls, const uint32_t* slot,
                          const uint64_t* off, uint8_t* out, hipStream_t st) {
  if (!R) return hipSuccess;
  hipLaunchKernelGGL(fa_copy_kernel, dim3((R + 3) / 4), dim3(256), 0, st, d, rec, R, cls, slot, off, out);
  return hipGetLastError();
}

hipError_t launch_fa_names(const uint8_t* d, const FaRecord* rec, uint32_t R, uint32_t cls, const uint32_t* slot,
                           const uint64_t* noff, uint8_t* names, hipStream_t st) {
  if (!R) return hipSuccess;
  hipLaunchKernelGGL(fa_names_kernel, dim3((R + 255) / 256), dim3(256), 0, st, d, rec, R, cls, slot, noff, names);
  return hipGetLastError();
}

hipError_t launch_nl_count(const uint8_t* d, uint64_t len, uint32_t* count, hipStream_t st) {
  const uint64_t nb = ig_blocks(len);
  if (!nb) return hipSuccess;
  hipLaunchKernelGGL(nl_count_kernel, dim3((uint32_t)nb), dim3(kIgThreads), 0, st, d, len, count);
  return hipGetLastError();
}

hipError_t launch_nl_write(const uint8_t* d, uint64_t len, const uint64_t* off, uint64_t* nlpos, hipStream_t st) {
  const uint64_t nb = ig_blocks(len);
  if (!nb) return hipSuccess;
  hipLaunchKernelGGL(nl_write_kernel, dim3((uint32_t)nb), dim3(kIgThreads), 0, st, d, len, off, nlpos);
  return hipGetLastError();
}

hipError_t launch_depth_rows(const uint8_t* d, uint64_t len, const uint64_t* nlpos, uint64_t n_nl, uint64_t n_lines,
                             uint32_t S, uint32_t cv_ext, uint64_t min_depth_len, DepthRow* rows, float* means,
                             uint32_t* flags, hipStream_t st) {
  if (!n_lines) return hipSuccess;
  hipLaunchKernelGGL(depth_rows_kernel, dim3((uint32_t)((n_lines + 255) / 256)), dim3(256), 0, st, d, len, nlpos,
                     n_nl, n_lines, S, cv_ext, min_depth_len, rows, means, flags);
  return hipGetLastError();
}

hipError_t launch_depth_insert(const uint8_t* d, const DepthRow* rows, uint64_t n_lines, unsigned long long* keys,
                               uint32_t* vals, uint64_t mask, hipStream_t st) {
  if (!n_lines) return hipSuccess;
  hipLaunchKernelGGL(depth_insert_kernel, dim3((uint32_t)((n_lines + 255) / 256)), dim3(256), 0, st, d, rows,
                     n_lines, keys, vals, mask);
  return hipGetLastError();
}

hipError_t launch_depth_lookup(const uint8_t* d, const DepthRow* rows, const float* means, uint32_t S,
                               const unsigned long long* keys, const uint32_t* vals, uint64_t mask,
                               const uint8_t* names, const uint64_t* noff, uint32_t n, float* out, uint32_t* flags,
                               uint32_t* first_missing, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(depth_lookup_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d, rows, means, S, keys, vals,
                     mask, names, noff, n, out, flags, first_missing);
  return hipGetLastError();
}

hipError_t launch_fill_ig(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65535);
  hipLaunchKernelGGL(fill_u32_ig_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace rbg
