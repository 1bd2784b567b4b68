// HIP kernels for gfx950 (MI355X, CDNA4, wave64).
//
// Data layout in HBM (all row-major, one row per contig):
//   seqs     contig bytes back to back, offsets[n+1]
//   wx       n x mpad  u8   winning element (0..135) per register, 0xFF unset
//   planes   n x 8W    u32  register classes bit-sliced: word (w, p) holds bit
//                           p of the class byte of registers 32w..32w+31, at
//                           plane_word(w, p): planes 0-3 of every group w,
//                           then planes 4-7 (a 16-byte quarter group per
//                           (w, half), so eight lanes reading eight groups'
//                           same half read one contiguous 128-byte line).
//                           Unset registers and the pad registers >= m hold
//                           0xFF; a class is < 136, so 0xFF never equals one.
//   unit     n x 16NC  f32  centred, L2-normalised rank rows, zero padded
//
// Register-match count of a pair (row a, column b):
//   cnt = sum_w popcount( ~OR_p (a[w][p] ^ b'[w][p]) )
// where b' is b with plane 0 cleared at b's unset/pad registers, so an unset
// register shows 0xFF on the row side and 0xFE on the column side and never
// counts.  A class byte is equal exactly when the reference's 16-bit folded
// winners are equal and non-zero (src/rabbitbin.cpp:1261-1293), because the
// class of (element x, register k) is the smallest element with the same
// 16-bit fold at k.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "rbg_kernels.h"

namespace rbg {
namespace {

constexpr int kElemCount = 136;

__device__ __forceinline__ uint32_t base_code(uint32_t c) {
  const uint32_t u = c & 0xDFu;  // fold lower case; only a/A..t/T map onto A..T
  return u == 'A' ? 0u : u == 'C' ? 1u : u == 'G' ? 2u : u == 'T' ? 3u : 255u;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Position of word (group w, plane p) in a row of 8W plane words.
__host__ __device__ constexpr uint32_t plane_word(uint32_t w, uint32_t p, uint32_t W) {
  return (p >> 2) * 4u * W + 4u * w + (p & 3u);
}

// The eight plane words of group w of a row.
__device__ __forceinline__ void load_word(const uint32_t* __restrict__ row, uint32_t w, uint32_t W,
                                          uint32_t* dst) {
#pragma unroll
  for (int p = 0; p < 8; ++p) dst[p] = row[plane_word(w, p, W)];
}

// Column-side transform of one 32-register word group, in place: clear plane
// 0 where the column register is unset (all eight bits set).
__device__ __forceinline__ void column_fix(uint32_t* b) {
  uint32_t all = b[0];
#pragma unroll
  for (int p = 1; p < 8; ++p) all &= b[p];
  b[0] &= ~all;
}

// (a ^ b) | c in one v_bitop3_b32 (truth table 0xDE over (S0, S1, S2) =
// (a, c, b): 0xDE = (0xF0 ^ 0xAA) | 0xCC).
__device__ __forceinline__ uint32_t xor_or(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, c, b, 0xDE);
}

// Registers of a 32-register group whose classes differ: OR over the eight
// planes of a ^ b, as one chain (8 instructions).
__device__ __forceinline__ uint32_t word_mismatch(const uint32_t* a, const uint32_t* b) {
  uint32_t mm = a[0] ^ b[0];
#pragma unroll
  for (int p = 1; p < 8; ++p) mm = xor_or(a[p], b[p], mm);
  return mm;
}

__device__ __forceinline__ uint32_t pair_count_rows(const uint32_t* __restrict__ ra,
                                                    const uint32_t* __restrict__ rb,
                                                    uint32_t W) {
  uint32_t mism = 0;
  for (uint32_t w = 0; w < W; ++w) {
    uint32_t a[8], b[8];
    load_word(ra, w, W, a);
    load_word(rb, w, W, b);
    column_fix(b);
    mism += __popc(word_mismatch(a, b));
  }
  return 32u * W - mism;
}

// ---------------------------------------------------------------------------
// Stage (a): sketch.  One workgroup of four waves sketches four contigs.
//
// Phase 1: 4-mer histograms in LDS (src/rabbitbin.cpp:1459-1484).  The
// 16-byte aligned chunks of the four contigs are numbered one after another
// and the workgroup's 256 threads stride over all of them, so the
// workgroup's time follows the sum of its contigs' lengths, not the longest
// (contig lengths are log-normal).  The four waves share one histogram per
// contig.  A thread takes a chunk plus the dword before it (the 3-base
// overlap of 4-mers that end in the chunk) and decodes them four bytes at a
// time:
//   the 2-bit code is ((byte >> 1) ^ (byte >> 2)) & 3 (A 0, C 1, G 2, T 3,
//   either case); a byte is a base when byte & 0xDF is the upper-case
//   letter of its code (any other byte breaks every 4-mer covering it, as
//   the reference's window reset does); bytes outside the contig count as
//   non-bases.
// A 4-mer ending at byte r is valid when bytes r-3..r all are.  Its window E
// holds codes r-3..r in bits 0..7 (earliest base lowest) and is counted as
// it is: the reference's forward code F is E with its 2-bit groups reversed
// and the reverse complement's code is 255 - E, so the windows whose
// canonical code min(F, RC) is v are exactly those with E = rev(v) or
// E = 255 - v (one slot when v is its own reverse complement), and the
// canonical count is read from those two bins when the weights are formed.
// Then the GC-normalised weights in FP64, in the reference's operation order
// (src/rabbitbin.cpp:1491-1516), kept as 1/w per element; the base
// probabilities and 1/total are computed once per contig.
//
// Phase 2: wave v owns registers k = 64v + l + 256s; for each register the
// winner is argmin_x fl((1/w_x) * T[x][k]) over present elements, ties to the
// smaller x (the sketch's strict-less update, src/probmh.cpp:417-431).
// Elements are visited in ascending T[x][k] (sT/sX, shared by the four
// contigs).  With wmin the smallest 1/w of a contig, every element at or
// after rank r has fl((1/w_x) * T_x) >= fl(wmin * T_r) (both factors are no
// smaller, and rounding is monotone), so once fl(wmin * T_r) > best no later
// element can win or tie and the walk stops: typically after a few ranks.
// Ranks are loaded and applied in groups of four with the stop tested per
// group (applying elements past the stop changes nothing).  A lane walks one
// register at a time and takes the wave's next register when its walk stops.
// The winners go to LDS; then a wave's 64 registers give each contig two
// plane-word groups: the eight ballots of a contig are gathered into lanes
// 0..15 by selects and written by one store (two 32-byte segments).
//
// The sequence buffer must be readable 16 bytes past its last contig (the
// last chunk of a contig is read whole).  A thread reads the dword before
// its chunk; for the first chunk of a workgroup it reads the chunk's first
// dword instead (those bytes stand before every contig of the workgroup, so
// no window or base takes them).
// ---------------------------------------------------------------------------
constexpr int kSketchWaves = 4;
constexpr int kRankGroup = 4;  // proposal ranks loaded together in phase 2

// Bit 7 of each byte set iff that byte of x is non-zero.
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t x) {
  return (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}
// One word of sequence: returns the 2-bit codes of its bytes (A, C, G, T =
// 0..3 in either case, read from bits 1-3 of the byte; byte b in bits
// 2b..2b+1) and sets x = the word XOR the upper-case letters of those codes
// (v_perm_b32 maps a code back to its letter), so byte b is one of A, C, G, T
// exactly when byte b of x & 0xDF is zero.
__device__ __forceinline__ uint32_t word_codes(uint32_t w, uint32_t& x) {
  const uint32_t c = ((w >> 1) ^ (w >> 2)) & 0x03030303u;
  const uint32_t letters = 0x54474341u;  // 'T' 'G' 'C' 'A'
  x = w ^ __builtin_amdgcn_perm(letters, letters, c);
  const uint32_t t = (c | (c >> 6)) & 0x000F000Fu;
  return (t | (t >> 12)) & 0xFFu;
}
// Bit b set iff byte b of the word is a base (from word_codes' x).
__device__ __forceinline__ uint32_t word_valid4(uint32_t x) {
  const uint32_t ok = ~nonzero_bytes(x & 0xDFDFDFDFu) & 0x80808080u;
  return ((ok >> 7) | (ok >> 14) | (ok >> 21) | (ok >> 28)) & 0xFu;
}

// Bit i of a 16-bit mask moved to bit 2i.
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  return (x | (x << 1)) & 0x55555555u;
}

struct KmerTally {
  uint32_t base[4] = {0, 0, 0, 0};
  uint32_t kmers = 0;
};

// Histogram slot E of the 4-mer ending at chunk byte i.  c40 holds the codes
// of the chunk's 20 stream bytes (the dword before the chunk, then its 16
// bytes; code t at bits 2t..2t+1), so the window of stream bytes i+1..i+4 is
// bits 2(i+1)..2(i+1)+7.
__device__ __forceinline__ uint32_t window_slot(uint64_t c40, uint32_t i) {
  return (uint32_t)(c40 >> (2 * (i + 1))) & 0xFFu;
}

// The 4-mers ending in the 16 bytes [16j, 16j+16), counted in pairs: the
// 5-mer ending at chunk byte 2t+1 (t = 0..7) holds the 4-mers ending at 2t
// (its first four bases) and 2t+1 (its last four), so one add to a 5-mer bin
// counts two windows.  Every lane counts all eight 5-mers (nearly every chunk
// is all bases and inside its contig); a lane whose chunk is not takes each
// 5-mer with an invalid window off again and adds its valid window, if any,
// to the 4-mer bins h4 (one per element: elem maps a window value E to the
// element of its canonical 4-mer).  prev is the dword at 16j-4, `first` the position
// 16j-4 and [beg, end) the contig, both relative to the workgroup's first
// chunk (a workgroup's contigs span less than 2^31 bytes).
__device__ __forceinline__ void chunk_kmers(uint32_t* __restrict__ h5, uint32_t* __restrict__ h4,
                                            const uint8_t* __restrict__ elem,
                                            uint32_t prev, uint4 cur, int32_t first, int32_t beg,
                                            int32_t end, KmerTally& tl) {
  uint32_t x[5];
  const uint32_t c0 = word_codes(prev, x[0]), c1 = word_codes(cur.x, x[1]),
                 c2 = word_codes(cur.y, x[2]), c3 = word_codes(cur.z, x[3]),
                 c4 = word_codes(cur.w, x[4]);
  const uint64_t c40 = (uint64_t)(c0 | (c1 << 8) | (c2 << 16) | (c3 << 24)) | ((uint64_t)c4 << 32);
  // 5-mer t: stream bytes 2t+1 .. 2t+5, bits 4t+2 .. 4t+11 of c40
  const uint32_t clo = (uint32_t)c40, chi = (uint32_t)(c40 >> 32);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const uint32_t v = t == 0 ? clo : __builtin_amdgcn_alignbit(chi, clo, 4 * t);
    atomicAdd(&h5[(v >> 2) & 0x3FFu], 1u);
  }
  const bool inside = first >= beg && first + 20 <= end;
  const bool full = inside && ((x[0] | x[1] | x[2] | x[3] | x[4]) & 0xDFDFDFDFu) == 0;
  uint32_t valid = 0x55555555u;  // bases among the 16 chunk bytes (2-bit spaced)
  uint32_t km = 0xFFFFu;         // valid windows
  if (!full) {
    uint32_t vm = word_valid4(x[0]) | (word_valid4(x[1]) << 4) | (word_valid4(x[2]) << 8) |
                  (word_valid4(x[3]) << 12) | (word_valid4(x[4]) << 16);
    const uint32_t lo_c = (uint32_t)min(max(beg - first, 0), 20);
    const uint32_t hi_c = (uint32_t)min(max(end - first, 0), 20);
    vm &= ((1u << hi_c) - 1u) & ~((1u << lo_c) - 1u);
    // bit i: the 4-mer ending at chunk byte i (buffer bytes 16j+i-3 .. 16j+i) is valid
    km = ((vm & (vm >> 1) & (vm >> 2) & (vm >> 3)) >> 1) & 0xFFFFu;
    const uint32_t inv = ~km & 0xFFFFu;
    // bit 2t: 5-mer t covers an invalid window
    for (uint32_t pm = (inv | (inv >> 1)) & 0x5555u; pm != 0; pm &= pm - 1u) {
      const uint32_t i = (uint32_t)__ffs(pm) - 1u;  // = 2t
      atomicSub(&h5[(uint32_t)(c40 >> (2 * i + 2)) & 0x3FFu], 1u);
      if ((km >> i) & 1u) atomicAdd(&h4[elem[window_slot(c40, i)]], 1u);
      if ((km >> (i + 1)) & 1u) atomicAdd(&h4[elem[window_slot(c40, i + 1)]], 1u);
    }
    valid = spread16((vm >> 4) & 0xFFFFu);
  }
  tl.kmers += __popc(km);
  const uint32_t c32 = (uint32_t)(c40 >> 8);  // codes of the 16 chunk bytes
#pragma unroll
  for (uint32_t b = 0; b < 4; ++b) {
    const uint32_t eq = ~(c32 ^ (0x55555555u * b));
    tl.base[b] += __popc(eq & (eq >> 1) & valid);
  }
}

// Windows with value E (earliest base lowest) in a contig counted by the
// 5-mer bins: those whose first four bases (E | x << 8) or last four
// (x | E << 2) are E.
__device__ __forceinline__ uint32_t window_count(const uint32_t* __restrict__ h5, uint32_t e) {
  uint32_t c = 0;
#pragma unroll
  for (uint32_t x = 0; x < 4; ++x) c += h5[e | (x << 8)] + h5[x | (e << 2)];
  return c;
}

// Plane words of a wave's 64 class bytes (one per lane) into lanes 0..15 of
// val: plane P's ballot gives group g0's word (lanes 0-31) at lane
// 8(P/4) + P%4 and group g0+1's (lanes 32-63) at that lane + 4.  Each lane
// picks its word from the wave-uniform ballots with selects.
template <int P>
__device__ __forceinline__ void plane_lanes(uint32_t& val, uint32_t code, uint32_t ln) {
  const uint64_t bits = __ballot((code >> P) & 1u);
  const bool mine = (ln & 3u) == (uint32_t)(P & 3) && ((ln >> 3) & 1u) == (uint32_t)(P >> 2);
  const uint32_t word = (ln & 4u) ? (uint32_t)(bits >> 32) : (uint32_t)bits;
  val = mine ? word : val;
  if constexpr (P < 7) plane_lanes<P + 1>(val, code, ln);
}

// A 4-mer code with its four 2-bit groups in reverse order.
__device__ __forceinline__ uint32_t rev4(uint32_t v) {
  return ((v & 3u) << 6) | ((v & 0xCu) << 2) | ((v >> 2) & 0xCu) | ((v >> 6) & 3u);
}

template <int CPB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void sketch_kernel(
    const uint8_t* __restrict__ seqs, const uint64_t* __restrict__ offs, uint32_t n, uint32_t m,
    uint32_t W, const double* __restrict__ sT, const uint8_t* __restrict__ sX,
    const uint8_t* __restrict__ cls, const uint8_t* __restrict__ codes, uint8_t* __restrict__ wx,
    uint32_t mpad, uint32_t* __restrict__ planes, const uint32_t* __restrict__ out_row) {
  static_assert(CPB == kSketchWaves, "four contigs per workgroup");
  // Phase-1 bins per contig: 1024 5-mer bins, then one bin per element for
  // the windows counted alone (next to a non-base or a contig end).  After
  // the counts are read the same bytes hold winv (1/w per element).  At
  // 19.2 KB of LDS per workgroup, eight workgroups fit a CU.
  constexpr uint32_t kBins = 1024 + kElemCount;
  __shared__ __align__(16) uint32_t hbuf[CPB * kBins];
  double (*winv)[kElemCount] = reinterpret_cast<double (*)[kElemCount]>(hbuf);
  static_assert(sizeof(double) * CPB * kElemCount <= sizeof(hbuf), "winv fits the bins");
  __shared__ uint32_t bases[CPB][4];
  __shared__ uint32_t total[CPB];
  // Phase-1 tallies per contig (windows, then bases A, C, G, T) in eight
  // copies: lane l adds to copy l % 8, so the lanes of a wave that add to one
  // counter hit eight words in eight banks instead of one word.
  constexpr uint32_t kTallyCopies = 8;
  __shared__ uint32_t tallies[CPB][5][kTallyCopies];
  __shared__ double pbase[CPB][4];     // per contig: base probabilities
  __shared__ double tinv[CPB];         // per contig: 1 / total
  __shared__ unsigned long long wmin_bits[CPB];  // positive doubles order as integers
  __shared__ uint64_t soff[CPB + 1];   // contig q is [soff[q], soff[q+1])
  __shared__ uint8_t elem[256];        // window value E -> element of its canonical 4-mer
  const uint32_t tid = threadIdx.x;
  const uint32_t ln = tid & 63u;
  const uint32_t c0 = blockIdx.x * CPB;

  for (uint32_t v = tid; v < CPB * kBins / 4; v += 256) reinterpret_cast<uint4*>(hbuf)[v] = make_uint4(0, 0, 0, 0);
  if (tid < CPB) {
    wmin_bits[tid] = 0x7FF0000000000000ull;  // +inf: no element present
  }
  if (tid < CPB * 5 * kTallyCopies) (&tallies[0][0][0])[tid] = 0;
  if (tid <= CPB) soff[tid] = offs[min(c0 + tid, n)];  // contigs past n are empty
  if (tid < kElemCount) {
    const uint32_t v = codes[tid];
    elem[rev4(v)] = (uint8_t)tid;
    elem[255u - v] = (uint8_t)tid;
  }
  __syncthreads();
  // Chunks are numbered from the workgroup's first chunk (the one holding its
  // first contig's first byte, at buffer byte 16 * (soff[0] / 16)), with
  // 32-bit offsets from there.  Contig q owns the chunks counted cpre[q] ..
  // cpre[q+1] - 1 in the workgroup; count g of contig q is chunk g + cdj[q],
  // and its bytes are [cbe[q].x, cbe[q].y).  A chunk holding the end of one
  // contig and the start of the next is counted once for each.
  __shared__ int2 cbe[CPB];
  __shared__ int32_t cdj[CPB];
  __shared__ uint32_t cpre[CPB + 1];
  if (tid <= CPB) {
    const int64_t c16 = (int64_t)(soff[0] >> 4);
    uint32_t acc = 0;
    for (uint32_t q = 0; q < tid; ++q) {
      const int64_t b = (int64_t)soff[q], e = (int64_t)soff[q + 1];
      acc += e > b ? (uint32_t)(((e - 1) >> 4) - (b >> 4) + 1) : 0u;
    }
    cpre[tid] = acc;
    if (tid < CPB) {
      cbe[tid] = make_int2((int32_t)((int64_t)soff[tid] - 16 * c16), (int32_t)((int64_t)soff[tid + 1] - 16 * c16));
      cdj[tid] = (int32_t)(((int64_t)soff[tid] >> 4) - c16) - (int32_t)acc;
    }
  }
  __syncthreads();
  {
    // workgroup-uniform values, held in scalar registers
    const uint32_t nall = (uint32_t)__builtin_amdgcn_readfirstlane((int)cpre[CPB]);
    const uint32_t p1 = (uint32_t)__builtin_amdgcn_readfirstlane((int)cpre[1]);
    const uint32_t p2 = (uint32_t)__builtin_amdgcn_readfirstlane((int)cpre[2]);
    const uint32_t p3 = (uint32_t)__builtin_amdgcn_readfirstlane((int)cpre[3]);
    const int32_t d1 = __builtin_amdgcn_readfirstlane(cdj[1]);
    const int32_t d2 = __builtin_amdgcn_readfirstlane(cdj[2]);
    const int32_t d3 = __builtin_amdgcn_readfirstlane(cdj[3]);
    const uint64_t blk_at = soff[0] & ~15ull;
    const uint8_t* blk = seqs + (((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(blk_at >> 32)) << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)blk_at));
    // contig and chunk of count g; the chunk and the dword before it (for the
    // workgroup's first chunk its own first dword: those bytes stand before
    // every contig of the workgroup, so no window or base takes them)
    auto locate = [p1, p2, p3, d1, d2, d3](uint32_t g, uint32_t& q, int32_t& jr) {
      const bool a1 = g >= p1, a2 = g >= p2, a3 = g >= p3;
      q = (uint32_t)a1 + (uint32_t)a2 + (uint32_t)a3;
      int32_t d = a1 ? d1 : 0;  // selects of values (a select of references becomes a load)
      d = a2 ? d2 : d;
      d = a3 ? d3 : d;
      jr = (int32_t)g + d;
    };
    auto load = [&](int32_t jr, uint4& cur, uint32_t& prev) {
      const uint32_t at = 16u * (uint32_t)jr;
      cur = *reinterpret_cast<const uint4*>(blk + at);
      prev = *reinterpret_cast<const uint32_t*>(blk + (max(at, 4u) - 4u));
    };
    KmerTally tl;
    uint32_t tq = 0;  // contig of the tally (a thread's chunks come in contig order)
    auto flush = [&]() {
      if (tl.kmers | tl.base[0] | tl.base[1] | tl.base[2] | tl.base[3]) {
        atomicAdd(&tallies[tq][0][ln % kTallyCopies], tl.kmers);
#pragma unroll
        for (int b = 0; b < 4; ++b) atomicAdd(&tallies[tq][1 + b][ln % kTallyCopies], tl.base[b]);
      }
    };
    // count one chunk while the thread's next one (count g + 256; past the end
    // the last chunk again, unused) is loaded
    auto step = [&](uint32_t g, uint32_t q, int32_t jr, uint4 cur, uint32_t prev, uint32_t& qn,
                    int32_t& jn, uint4& nxt, uint32_t& nprev) {
      locate(min(g + 256u, nall - 1u), qn, jn);
      load(jn, nxt, nprev);
      if (q != tq) {  // the tally so far belongs to contig tq
        flush();
        tl = KmerTally();
        tq = q;
      }
      const int2 be = cbe[q];
      chunk_kmers(hbuf + q * 1024u, hbuf + CPB * 1024u + q * kElemCount, elem, prev, cur, 16 * jr - 4, be.x, be.y, tl);
    };
    uint32_t g = tid;
    if (g < nall) {
      uint32_t qa, qb;
      int32_t ja, jb;
      uint4 ca, cb;
      uint32_t pa, pb;
      locate(g, qa, ja);
      load(ja, ca, pa);
      tq = qa;
      // two chunk buffers used in turn, so that no registers are copied
      while (true) {
        step(g, qa, ja, ca, pa, qb, jb, cb, pb);
        g += 256;
        if (g >= nall) break;
        step(g, qb, jb, cb, pb, qa, ja, ca, pa);
        g += 256;
        if (g >= nall) break;
      }
      flush();
    }
  }
  __syncthreads();
  if (tid < CPB) {  // per-contig factors, in the reference's operations
    const uint32_t q = tid;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t r = 0; r < kTallyCopies; ++r) sum += tallies[q][v][r];
      if (v == 0)
        total[q] = sum;
      else
        bases[q][v - 1] = sum;
    }
    const uint64_t vb = (uint64_t)bases[q][0] + bases[q][1] + bases[q][2] + bases[q][3];
#pragma unroll
    for (int b = 0; b < 4; ++b) pbase[q][b] = vb > 0 ? (double)bases[q][b] / (double)vb : 0.25;
    tinv[q] = total[q] != 0 ? 1.0 / (double)total[q] : 0.0;
  }
  // canonical counts of this thread's elements e = tid + 256r: code v's
  // windows have E = rev(v) or E = 255 - v (one value when v is its own
  // reverse complement)
  constexpr int kEpt = (CPB * kElemCount + 255) / 256;
  uint32_t ecnt[kEpt];
#pragma unroll
  for (int r = 0; r < kEpt; ++r) {
    const uint32_t e = tid + 256u * r;
    ecnt[r] = 0;
    if (e < CPB * kElemCount) {
      const uint32_t q = e / kElemCount, x = e % kElemCount;
      const uint32_t v = codes[x];
      const uint32_t sa = rev4(v), sb = 255u - v;
      const uint32_t* h5 = hbuf + q * 1024u;
      ecnt[r] = hbuf[CPB * 1024u + q * kElemCount + x] + window_count(h5, sa) + (sa != sb ? window_count(h5, sb) : 0u);
    }
  }
  __syncthreads();  // bins read: their bytes now take winv
#pragma unroll
  for (int r = 0; r < kEpt; ++r) {
    const uint32_t e = tid + 256u * r;
    if (e >= CPB * kElemCount) break;
    const uint32_t q = e / kElemCount, x = e % kElemCount;
    const uint32_t c = c0 + q;
    const uint32_t v = codes[x];
    const uint32_t cnt = ecnt[r];
    double wi = __builtin_huge_val();  // absent element: fl(inf * T) = inf never wins
    if (c < n && cnt != 0 && total[q] != 0) {
      double pf = 1.0, pr = 1.0;
      uint32_t code = v;
#pragma unroll
      for (int pos = 0; pos < 4; ++pos) {
        const uint32_t b = code & 3u;
        pf *= pbase[q][b];
        pr *= pbase[q][3u - b];
        code >>= 2;
      }
      const double obs = (double)cnt * tinv[q];
      const double ex = pf + pr;
      double w = 0.0;
      if (ex > 1e-14) {
        w = obs / ex;
        if (w > 20.0) w = 20.0;
      }
      if (w > 0.0) wi = 1.0 / w;
    }
    winv[q][x] = wi;
    if (wi != __builtin_huge_val())
      atomicMin(&wmin_bits[q], (unsigned long long)__double_as_longlong(wi));
  }
  __syncthreads();
  double wmin[CPB];
  uint32_t orow[CPB];  // output row of each contig
#pragma unroll
  for (int q = 0; q < CPB; ++q) {
    // uniform: kept in scalar registers
    const unsigned long long wb = wmin_bits[q];
    const uint32_t wlo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)wb);
    const uint32_t whi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(wb >> 32));
    wmin[q] = __longlong_as_double((long long)(((uint64_t)whi << 32) | wlo));
    const uint32_t c = c0 + q;
    orow[q] = c < n ? (out_row ? out_row[c] : c) : 0u;
  }

  // Registers k = 64v + l + 256s (s < kspan / 256) belong to wave v.  A lane
  // walks one register at a time; when its walk stops it records the four
  // winners in wbx[k] (byte q: contig q's winner, 0xFF for none) and takes
  // the wave's next unwalked register, so that the wave's lanes stay busy
  // until its registers run out (walk lengths vary: about 1.8 rank groups per
  // register on average, but the longest of 64 registers is several times
  // that).  The winners are turned into output bytes and plane words
  // afterwards, with lane l of wave v on registers 64v + l + 256s again.
  const uint32_t kspan = ((32u * W + 255u) / 256u) * 256u;  // <= 512 (m <= 512)
  uint32_t* wbx = hbuf + 2 * CPB * kElemCount;                 // after winv
  static_assert(2 * CPB * kElemCount + 512 <= CPB * kBins, "winners fit the bins");
  static_assert(kElemCount % kRankGroup == 0, "whole rank groups");
  {
    const uint32_t wv = tid >> 6;
    const uint32_t npool = kspan / 4u;  // registers per wave
    auto kof = [&](uint32_t p) { return 64u * wv + 256u * (p >> 6) + (p & 63u); };
    uint32_t p = ln, pn = 64;  // this lane's pool entry, the wave's next one
    uint32_t k = kof(p);
    bool act = p < npool && k < m;  // kof grows with p: once k >= m, every later k too
    int r0 = 0;
    uint32_t at = k;  // table entry of rank r0 at register k (< 136 * 512)
    double best[CPB];
    int bx[CPB];
#pragma unroll
    for (int q = 0; q < CPB; ++q) {
      best[q] = __builtin_huge_val();
      bx[q] = -1;
    }
    while (__ballot(act)) {
      bool fin = false;
      if (act) {
        // Ranks in groups of kRankGroup: applying an element past the break
        // point changes nothing (it can neither win nor tie), so the break
        // is tested once per group, on the group's last rank.
        double t[kRankGroup];
        uint32_t x[kRankGroup];
#pragma unroll
        for (int g = 0; g < kRankGroup; ++g) {  // rank r0 + g at element (r0 + g) * m + k
          const uint32_t e = at + (uint32_t)g * m;
          t[g] = *reinterpret_cast<const double*>(reinterpret_cast<const char*>(sT) + 8u * e);
          x[g] = sX[e];
        }
        // An absent element has 1/w = inf, hence hv = inf, which neither
        // beats nor ties best (best = inf only while bx = -1, and x > -1).
#pragma unroll
        for (int g = 0; g < kRankGroup; ++g) {
          double wq[CPB];
#pragma unroll
          for (int q = 0; q < CPB; ++q) wq[q] = winv[q][x[g]];
#pragma unroll
          for (int q = 0; q < CPB; ++q) {
            const double hv = wq[q] * t[g];
            const bool take = (hv < best[q]) | ((hv == best[q]) & ((int)x[g] < bx[q]));
            best[q] = take ? hv : best[q];
            bx[q] = take ? (int)x[g] : bx[q];
          }
        }
        bool done = true;
#pragma unroll
        for (int q = 0; q < CPB; ++q)  // wmin = +inf: contig without elements
          done = done && (wmin[q] == __builtin_huge_val() || wmin[q] * t[kRankGroup - 1] > best[q]);
        fin = done || r0 + kRankGroup >= kElemCount;
        if (fin) {
          uint32_t packed = 0;
#pragma unroll
          for (int q = 0; q < CPB; ++q) packed |= ((uint32_t)bx[q] & 0xFFu) << (8 * q);
          wbx[k] = packed;
        }
        r0 += kRankGroup;
        at += kRankGroup * m;
      }
      const uint64_t fm = __ballot(fin);
      if (fin) {  // the next registers, in lane order
        p = pn + lanes_below(fm);
        k = kof(p);
        act = p < npool && k < m;
        r0 = 0;
        at = k;
#pragma unroll
        for (int q = 0; q < CPB; ++q) {
          best[q] = __builtin_huge_val();
          bx[q] = -1;
        }
      }
      pn += (uint32_t)__popcll(fm);
    }
  }
  __syncthreads();  // winners recorded (a register's walker is any lane of its wave)
  for (uint32_t k = tid; k < kspan; k += 256) {
    const uint32_t wb = k < m ? wbx[k] : 0xFFFFFFFFu;
    int bx[CPB];
#pragma unroll
    for (int q = 0; q < CPB; ++q) {
      const uint32_t b = (wb >> (8 * q)) & 0xFFu;
      bx[q] = b == 0xFFu ? -1 : (int)b;
    }
    // class bytes of the four winners, loaded together
    uint32_t code[CPB];
#pragma unroll
    for (int q = 0; q < CPB; ++q) {
      const bool has = k < m && bx[q] >= 0;
      const uint32_t cv = cls[has ? (size_t)k * kElemCount + (uint32_t)bx[q] : 0];
      code[q] = has ? cv : 0xFFu;
    }
    if (wx != nullptr && k < mpad) {
#pragma unroll
      for (int q = 0; q < CPB; ++q)
        if (c0 + q < n) wx[(size_t)orow[q] * mpad + k] = (k < m && bx[q] >= 0) ? (uint8_t)bx[q] : (uint8_t)0xFF;
    }
    // plane words of the wave's 64 registers: groups g0 (lanes 0-31) and g0+1
    const uint32_t g0 = (k - ln) / 32u;
    const uint32_t s = ln & 7u, h = (ln >> 3) & 1u;  // lane < 16 writes word (group g0 + s/4, plane 4h + s%4)
#pragma unroll
    for (int q = 0; q < CPB; ++q) {
      uint32_t val = 0;
      plane_lanes<0>(val, code[q], ln);
      if (ln < 16 && c0 + q < n && g0 + (s >> 2) < W)
        planes[(size_t)orow[q] * 8u * W + h * 4u * W + 4u * g0 + s] = val;
    }
  }
}

// Bit planes from class bytes: one wave per (row, 64-register span); lane l
// holds register 64s + l, and each plane's ballot gives two plane words.
__global__ __launch_bounds__(256) void planes_from_classes_kernel(const uint8_t* __restrict__ cls,
                                                                  uint32_t n, uint32_t m, uint32_t mpad,
                                                                  uint32_t W, uint32_t row0,
                                                                  uint32_t* __restrict__ planes) {
  const uint32_t spans = (32u * W + 63u) / 64u;
  const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const uint32_t ln = threadIdx.x & 63u;
  if (wave >= (uint64_t)n * spans) return;  // wave-uniform
  const uint32_t r = (uint32_t)(wave / spans), sp = (uint32_t)(wave % spans);
  const uint32_t k = 64u * sp + ln;
  const uint32_t code = k < m ? cls[(size_t)r * mpad + k] : 0xFFu;
  const uint32_t w0 = 2u * sp + (ln >> 5);
  uint32_t* row = planes + (size_t)(row0 + r) * 8u * W;
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const uint64_t bits = __ballot((code >> p) & 1u);
    if ((ln == 0 || ln == 32) && w0 < W) row[plane_word(w0, p, W)] = ln == 0 ? (uint32_t)bits : (uint32_t)(bits >> 32);
  }
}

// Counts for an explicit list of pairs (b0 sampling).
__global__ __launch_bounds__(256) void list_counts_kernel(const uint32_t* __restrict__ planes,
                                                          uint32_t W, const uint32_t* __restrict__ pi,
                                                          const uint32_t* __restrict__ pj,
                                                          uint32_t np, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= np) return;
  out[t] = pair_count_rows(planes + (size_t)pi[t] * 8u * W, planes + (size_t)pj[t] * 8u * W, W);
}

// 16-bit winner compare for a pair list (src/rabbitbin.cpp:1261-1293).
__global__ __launch_bounds__(256) void win16_counts_kernel(const uint16_t* __restrict__ w16,
                                                           uint32_t m, const uint32_t* __restrict__ pi,
                                                           const uint32_t* __restrict__ pj,
                                                           uint32_t np, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t >= np) return;
  const uint16_t* a = w16 + (size_t)pi[t] * m;
  const uint16_t* b = w16 + (size_t)pj[t] * m;
  uint32_t c = 0;
  for (uint32_t k = 0; k < m; ++k) c += (a[k] != 0 && a[k] == b[k]) ? 1u : 0u;
  out[t] = c;
}

// ---------------------------------------------------------------------------
// Dense counts (the small regime, n <= 25000, and the fused regime without the
// abundance gate): every pair's count into an n x n u16 matrix, computed once
// per unordered pair.  Thread = column j with its bit planes in VGPRs; rows
// r < j are uniform across the wave and come in through scalar loads.  The
// count is symmetric (a register matches only where both sketches set it), so
// (r, j) is written at [r][j] (lanes write a coalesced row segment) and at
// [j][r]: the counts of 64 rows are staged in LDS, and each wave then writes
// 64 columns' row segments [j][r0 .. r0+64), one 128-byte segment per store
// instruction.  The diagonal is neither written nor read.
// ---------------------------------------------------------------------------
// Work item (assembly a, column block cb, row block rb, rows): thread = local
// column 256*cb + t of assembly a against its rows [rb*rows, +rows), pairs
// r < j with j >= jmin only (jmin > 0: a served session's earlier pairs are in
// the matrix already); items without such a pair are not generated.
constexpr uint32_t kDenseSub = 64;  // rows staged per LDS tile
template <int W>
__global__ __launch_bounds__(256) void dense_counts_kernel(const uint32_t* __restrict__ planes,
                                                           const AsmDesc* __restrict__ desc,
                                                           const uint4* __restrict__ items,
                                                           uint16_t* __restrict__ mat) {
  __shared__ uint16_t tile[kDenseSub][256 + 2];  // odd word stride: column reads spread over banks
  const uint4 it = items[blockIdx.x];
  const uint32_t n = desc[it.x].n, jmin = desc[it.x].jmin;
  const uint32_t* __restrict__ rows = planes + (size_t)desc[it.x].row0 * 8 * W;
  uint16_t* __restrict__ out = mat + desc[it.x].mat0;
  const uint32_t tid = threadIdx.x, ln = tid & 63u, wv = tid >> 6;
  const uint32_t j0 = it.y * 256;
  const uint32_t j = j0 + tid;
  const uint32_t jlast = min(n, j0 + 256) - 1;  // the block's last column
  const uint32_t r0 = it.z * it.w;
  const uint32_t r1 = min(min(n, r0 + it.w), jlast);  // rows with some column j > r
  uint32_t b[8 * W];
  const uint32_t jj = j < n ? j : n - 1;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    load_word(rows + (size_t)jj * 8 * W, w, W, b + 8 * w);
    column_fix(b + 8 * w);
  }
  const bool jok = j < n && j >= jmin;
  for (uint32_t rs = r0; rs < r1; rs += kDenseSub) {
    const uint32_t re = min(r1, rs + kDenseSub);
    for (uint32_t r = rs; r < re; ++r) {
      const uint32_t* __restrict__ ra = rows + (size_t)r * 8 * W;
      uint32_t mism = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        uint32_t a[8];
        load_word(ra, w, W, a);
        mism += __popc(word_mismatch(a, b + 8 * w));
      }
      const uint16_t c = (uint16_t)(32u * W - mism);
      tile[r - rs][tid] = c;
      if (jok && j > r) out[(size_t)r * n + j] = c;
    }
    __syncthreads();
    // [j][r] for this wave's 64 columns: lane = row r = rs + ln
    const uint32_t r = rs + ln;
    for (uint32_t q = 0; q < 64; ++q) {
      const uint32_t col = 64 * wv + q, jc = j0 + col;
      if (jc < n && jc >= jmin && r < re && jc > r) out[(size_t)jc * n + r] = tile[ln][col];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused regime, pair pass in two kernels:
//  1. gate_mfma_kernel: int8 MFMA prefilter of the abundance gate over the
//     upper triangle, emits candidates (i < j, qdot >= qcut);
//  2. gate_compare_kernel: is_nz and the exact FP32 gate per candidate, and
//     the register compare of the survivors, in place.
// qcut is a rigorous lower bound (gate_qcut), so the survivors are exactly
// the pairs the reference's gate keeps (src/impl/rb_graph.cpp:1108-1116).
//
// gate_mfma_kernel: work item = a 256-row stripe bi and up to 16 consecutive
// 256-column tiles bj0.. (bi <= bj).  Wave w owns rows bi*256+64w..+64 as two
// 32-row operands held in registers for the whole item; the tiles' column
// operands are double-buffered in LDS (the next tile's loads are in flight
// while the current tile computes).  v_mfma_i32_32x32x32_i8 (the gfx950
// double-rate form: 32 x 32 pairs at K = 32 per instruction) gives the integer
// dot of the quantised rows: lane l supplies row (or column) l%32, bytes
// 16*(l/32)..+15 of each 32-byte K chunk, and holds in register v the pair
// (row 8*(v/4) + 4*(l/32) + v%4, column l%32).  Per 32-column step a wave
// makes 2048 pair dots with 2*KC MFMAs and takes the lane maximum with 16
// max3; only a step whose maximum reaches qcut looks at single pairs.
// Candidates are staged per wave in LDS as local coordinates (row in the
// stripe << 16 | column from the item's first tile: items of up to 256 tiles;
// bit 31: the dot reaches qcut_hi, so the pair passes the exact gate) and
// appended to `cand` with one global atomic per flush (the proof as bit 15).
// ---------------------------------------------------------------------------
typedef int v4i32 __attribute__((ext_vector_type(4)));
typedef int v16i32 __attribute__((ext_vector_type(16)));
constexpr uint32_t kGateBuf = 512;  // per-wave candidate buffer, flushed when a ballot may not fit

// Lane slice of a q8 row (zero beyond n): bytes 32c + 16h .. +15 per K chunk c.
// The offset is 32-bit (an assembly's q8 rows span < 2^30 bytes), so the load
// takes the assembly's base in scalar registers.
template <int KC>
__device__ __forceinline__ void load_q8_slice(const int8_t* __restrict__ q, uint32_t row, uint32_t n,
                                              uint32_t h, v4i32* out) {
#pragma unroll
  for (int c = 0; c < KC; ++c)
    out[c] = row < n ? *reinterpret_cast<const v4i32*>(q + (row * (32u * KC) + 32u * c + 16u * h))
                     : v4i32{0, 0, 0, 0};
}

// This thread's 2 entries of a tile's column operands: entry e = tid + 256k
// is (step e/64, lane e%64), column 32*(e/64) + e%32, half (e%64)/32.
template <int KC>
__device__ __forceinline__ void load_tile_ops(const int8_t* __restrict__ q, uint32_t jb, uint32_t n,
                                              v4i32 (*dst)[KC]) {
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const uint32_t e = threadIdx.x + 256u * k;
    load_q8_slice<KC>(q, jb + 32u * (e >> 6) + (e & 31u), n, (e >> 5) & 1u, dst[k]);
  }
}

// Six waves per SIMD for S <= 32 (81 VGPRs unconstrained: five); the one
// spilled value is reloaded per tile before the tile's loads are issued.
template <int KC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KC == 1 ? 6 : 1))) void gate_mfma_kernel(const int8_t* __restrict__ q_all,
                                                        const AsmDesc* __restrict__ desc,
                                                        const uint32_t* __restrict__ perm_all,
                                                        const uint4* __restrict__ items,
                                                        uint64_t* __restrict__ cand,
                                                        unsigned long long* __restrict__ ncand,
                                                        uint64_t cap) {
  __shared__ uint32_t buf[4][kGateBuf];
  __shared__ v4i32 bops[2][8][KC][64];
  const uint32_t wv = threadIdx.x >> 6, ln = threadIdx.x & 63u;
  const uint4 item = items[blockIdx.x];
  // the item's assembly: its rows, in pair-pass order, start at row0
  const uint32_t n = desc[item.x].n;
  const int32_t qcut = desc[item.x].qcut, qcut_hi = desc[item.x].qcut_hi;
  const int8_t* __restrict__ q = q_all + (size_t)desc[item.x].row0 * 32 * KC;
  const uint32_t* __restrict__ perm = perm_all + desc[item.x].row0;
  const uint32_t ib = (item.y >> 16) * 256u, jbase = (item.y & 0xFFFFu) * 256u;
  const uint32_t ntile = item.z;
  const uint32_t jmin = item.w;  // pairs with j < jmin were evaluated by an earlier build
  // the wave's first row, in a scalar register: the step skips below are branches
  const uint32_t row0 = __builtin_amdgcn_readfirstlane(ib + 64 * wv);

  v4i32 A[2][KC];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) load_q8_slice<KC>(q, row0 + 32 * rb + (ln & 31u), n, ln >> 5, A[rb]);
  v4i32 nb[2][KC];
  load_tile_ops<KC>(q, jbase, n, nb);
  uint32_t bn = 0;  // wave-uniform count of staged candidates

  auto flush = [&]() {
    __builtin_amdgcn_wave_barrier();  // entries written by other lanes
    unsigned long long base = 0;
    if (ln == 0) base = atomicAdd(ncand, (unsigned long long)bn);
    base = __shfl(base, 0);
    for (uint32_t t = ln; t < bn; t += 64) {
      const uint32_t e = buf[wv][t];
      const uint64_t a = perm[ib + ((e >> 16) & 0x7FFFu)], b = perm[jbase + (e & 0xFFFFu)];
      const uint64_t i = a < b ? a : b, j = a < b ? b : a;
      if (base + t < cap) cand[base + t] = (i << 40) | (j << 16) | (uint64_t)((e >> 31) << 15);
    }
    bn = 0;
  };

#pragma unroll 1
  for (uint32_t t = 0; t < ntile; ++t) {
    // tile t's operands (prefetched) go to LDS (the buffer of tile t-2: every
    // wave passed the barrier of tile t-1 after reading it); then prefetch t+1
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const uint32_t e = threadIdx.x + 256u * k;
#pragma unroll
      for (int c = 0; c < KC; ++c) bops[t & 1][e >> 6][c][e & 63u] = nb[k][c];
    }
    __syncthreads();
    if (t + 1 < ntile) load_tile_ops<KC>(q, jbase + 256u * (t + 1), n, nb);
    const uint32_t jb = jbase + 256u * t;
    const bool diag = ib == jb;
#pragma unroll 2
    for (uint32_t s = 0; s < 8; ++s) {
      const uint32_t j0 = jb + 32u * s;
      if (j0 >= n) break;
      if (diag && j0 + 31 <= row0) continue;  // every pair of the step has j <= i
      v4i32 B[KC];
#pragma unroll
      for (int c = 0; c < KC; ++c) B[c] = bops[t & 1][s][c][ln];
      v16i32 acc0 = {}, acc1 = {};
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[0][c], B[c], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[1][c], B[c], acc1, 0, 0, 0);
      }
      int32_t mx = acc0[0];  // the largest of the lane's 32 dots (a chain of max3)
#pragma unroll
      for (int v = 1; v < 16; ++v) mx = max(mx, acc0[v]);
#pragma unroll
      for (int v = 0; v < 16; ++v) mx = max(mx, acc1[v]);
      if (__ballot(mx >= qcut) == 0) continue;  // no pair of these 2048 is near the gate
      // The lane's first row within the stripe, made opaque here so that the
      // compiler derives the row ids inside this rarely taken path instead of
      // keeping them in registers across the loop (which costs occupancy).
      uint32_t il0 = 64 * wv + 4 * (ln >> 5);
      asm volatile("" : "+v"(il0));
      const uint32_t j = j0 + (ln & 31u);
      const uint32_t jl = j - jbase;
      const bool full = !diag && jb + 256 <= n && jb >= jmin;
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const v16i32& acc = rb ? acc1 : acc0;
        int32_t mr = acc[0];
#pragma unroll
        for (int v = 1; v < 16; ++v) mr = max(mr, acc[v]);
        if (__ballot(mr >= qcut) == 0) continue;  // only the 32-row blocks with candidates
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const uint32_t il = il0 + 32 * rb + 8 * (v / 4) + (v % 4);  // row ib + il
          const bool c = full ? acc[v] >= qcut
                              : (acc[v] >= qcut && ib + il < j && j < n && j >= jmin);
          const uint64_t mask = __ballot(c);
          if (mask) {
            if (bn > kGateBuf - 64) flush();
            const uint32_t sure = acc[v] >= qcut_hi ? 1u : 0u;  // proven to pass the exact gate
            if (c) buf[wv][bn + lanes_below(mask)] = (sure << 31) | (il << 16) | jl;
            bn += (uint32_t)__popcll(mask);
          }
        }
      }
    }
  }
  if (bn) flush();
}

// Slots from a counter per key, atomicAdd(&ctr[key], 1) for every active
// lane, without one returning atomic per lane: the lanes are first grouped by
// key with ballots (no memory traffic), then every group's leader adds its
// group's size in one atomic instruction, so the wave waits for one round
// trip.  The first kGroups keys are grouped; the lanes left over form groups
// of one (their own leader).  Lanes with equal keys get consecutive slots:
//   g = wave_group_budget<8>(key, active);
//   if (g.count) base = atomicAdd(&ctr[key], g.count);
//   slot = wave_slot_finish(base, g);
// All lanes of the wave must call both.
struct WaveSlot {
  uint32_t leader = 0, rank = 0, count = 0;  // count: on leader lanes
};
template <int kGroups>
__device__ __forceinline__ WaveSlot wave_group_budget(uint32_t key, bool active) {
  const uint32_t ln = lane_id();
  WaveSlot g;
  g.leader = ln;
  g.count = active ? 1u : 0u;
  uint64_t pending = __ballot(active);
  for (int it = 0; pending != 0 && it < kGroups; ++it) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t k = __shfl(key, leader);
    const uint64_t grp = __ballot(active && key == k);
    if (active && key == k) {
      g.leader = (uint32_t)leader;
      g.rank = lanes_below(grp);
      g.count = (int)ln == leader ? (uint32_t)__popcll(grp) : 0u;
    }
    pending &= ~grp;
  }
  return g;
}
__device__ __forceinline__ uint32_t wave_slot_finish(uint32_t base, const WaveSlot& g) {
  return (uint32_t)__shfl(base, (int)g.leader) + g.rank;
}

// Gated pairs are packed as (i << 40) | (j << 16) | count; a candidate that
// the exact gate rejects is overwritten with kDeadPair.
constexpr uint64_t kDeadPair = ~0ull;

__device__ __forceinline__ void unpack_pair(uint64_t e, uint32_t& i, uint32_t& j, uint32_t& c) {
  i = (uint32_t)(e >> 40);
  j = (uint32_t)(e >> 16) & 0xFFFFFFu;
  c = (uint32_t)(e & 0xFFFFu);
}

// Exact gate and register compare of the int8 candidates, in place.
//
// Gate: keep (i < j) iff is_nz(i, j) and !(unit_dot_f(u_i, u_j) < thr), with
// unit_dot_f in the AVX-512 order of the reference build
// (src/rabbitbin.cpp:754-776): 16 accumulator lanes; chunk c (16 elements,
// zero padded) accumulates into acc1 when c is odd and lies inside a full
// 32-wide chunk (c < acc1_end), else into acc0; one FMA per chunk per lane
// in chunk order (padding adds fma(0,0,acc) = acc, as the masked tail does);
// then acc0+acc1 and GCC's _mm512_reduce_add_ps tree
// (l, l+8) -> (l, l+4) -> (0,2),(1,3) -> final.
//
// Eight GPU lanes take one candidate.  For the gate, lane q uses q4 = q & 3
// and holds accumulator lanes 4q4..4q4+3 (float4 q4 of every 16-float chunk
// of both rows; lanes q and q + 4 compute the same values); the first
// reduction level pairs lanes q and q^2, the second q and q^1, the last runs
// in every lane (IEEE addition is commutative, so which lane holds which
// operand does not change a sum, and every lane of the eight ends with the
// same dot).  The reductions move values between lanes with DPP quad
// permutes (and a half-row mirror for the register counts), not LDS.  For
// the register compare, lane q holds word groups q, q + 8, ... of both
// rows' planes (32 bytes each: one group's 8 planes, in two 16-byte halves),
// so each load instruction of the eight lanes of a candidate reads one
// contiguous 128-byte line of a row (the same half of eight groups).  The
// candidate e comes in from the previous step, and its plane rows are loaded
// before anything is used, so a step waits for one memory round trip: nearly
// every candidate survives.  The FP32 gate (unit rows, anynz, the assembly's
// parameters when the build holds several, the accumulation and reductions)
// runs only in steps where a candidate of the wave was not proven by the
// prefilter (about 6% of them at 300k contigs).
// Returns true on lane 0 of a survivor's eight, with its rows in ri, rj.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __uint_as_float(dpp_u32<CTRL>(__float_as_uint(v)));
}
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppHalfMirror = 0x141;  // lane i of each 8 from lane 7 - i

template <int NC, int W>
__device__ __forceinline__ bool gate_compare_one(
    uint64_t* __restrict__ cand, uint64_t t, uint64_t e, bool live, uint32_t q,
    const float* __restrict__ unit, const uint8_t* __restrict__ anynz,
    const AsmDesc* __restrict__ desc, const uint16_t* __restrict__ asm_of, bool single, float thr0,
    uint32_t acc1_end0, const uint32_t* __restrict__ planes, uint32_t& ri, uint32_t& rj) {
  constexpr int GPL = (W + 7) / 8;  // word groups per lane
  // Row indices are taken as 64-bit fields and masked in 64 bits: with the
  // 32-bit form hipcc (ROCm 7.2) dropped the 24-bit mask of j in the row
  // address of some instantiations while keeping it elsewhere.
  const uint64_t i64 = e >> 40, j64 = (e >> 16) & 0xFFFFFFull;
  const uint32_t i = (uint32_t)i64, j = (uint32_t)j64;
  // proven by the prefilter (its dot reached qcut_hi): no FP32 gate needed
  const bool sure = live && ((e >> 15) & 1u);
  const bool need = live && !sure;
  const uint32_t q4 = q & 3u;
  // The plane loads have no branch (a dead lane reads row 0's planes).
  uint4 pa[2 * GPL], pb[2 * GPL];
#pragma unroll
  for (int x = 0; x < GPL; ++x) {
    const uint32_t w = 8u * x + q;
    const uint32_t wl = w < (uint32_t)W ? w : 0u;  // lanes past the last group re-read group 0 (unused)
    const uint4* wa = reinterpret_cast<const uint4*>(planes + i64 * 8u * W + 4u * wl);
    const uint4* wb = reinterpret_cast<const uint4*>(planes + j64 * 8u * W + 4u * wl);
    pa[2 * x] = wa[0];             // planes 0-3 of group w
    pa[2 * x + 1] = wa[W];         // planes 4-7
    pb[2 * x] = wb[0];
    pb[2 * x + 1] = wb[W];
  }
  bool pass = sure;
  // The FP32 gate, only when a candidate of the wave needs it (nearly every
  // candidate is proven): its unit rows, is_nz inputs and the assembly's
  // parameters are read by the lanes that need them, and the arithmetic and
  // lane reductions run on the wave (zeros elsewhere).
  if (__ballot(need)) {  // wave-uniform
    float4 ua[NC], ub[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) ua[c] = ub[c] = float4{0.f, 0.f, 0.f, 0.f};
    uint32_t nzi = 1u, nzj = 1u;
    float thr = thr0;
    uint32_t acc1_end = acc1_end0;
    if (need) {
      const float4* ra = reinterpret_cast<const float4*>(unit + i64 * 16 * NC);
      const float4* rb = reinterpret_cast<const float4*>(unit + j64 * 16 * NC);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        ua[c] = ra[4 * c + q4];
        ub[c] = rb[4 * c + q4];
      }
      if (anynz) {
        nzi = anynz[i];
        nzj = anynz[j];
      }
      if (!single) {
        const AsmDesc& ad = desc[asm_of[i]];
        thr = ad.thr;
        acc1_end = ad.acc1_end;
      }
    }
    float acc0[4] = {0.f, 0.f, 0.f, 0.f}, acc1[4] = {0.f, 0.f, 0.f, 0.f};
    bool used1 = false;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float a4[4] = {ua[c].x, ua[c].y, ua[c].z, ua[c].w};
      const float b4[4] = {ub[c].x, ub[c].y, ub[c].z, ub[c].w};
      if ((c & 1) && (uint32_t)c < acc1_end) {
        used1 = true;
#pragma unroll
        for (int l = 0; l < 4; ++l) acc1[l] = __builtin_fmaf(a4[l], b4[l], acc1[l]);
      } else {
#pragma unroll
        for (int l = 0; l < 4; ++l) acc0[l] = __builtin_fmaf(a4[l], b4[l], acc0[l]);
      }
    }
    if (used1) {
#pragma unroll
      for (int l = 0; l < 4; ++l) acc0[l] = acc0[l] + acc1[l];
    }
    float t8[4], t4[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) t8[l] = dpp_f32<kDppXor2>(acc0[l]) + acc0[l];  // l+8 with l
#pragma unroll
    for (int l = 0; l < 4; ++l) t4[l] = dpp_f32<kDppXor1>(t8[l]) + t8[l];  // l+4 with l
    const float dot = (t4[0] + t4[2]) + (t4[1] + t4[3]);
    pass = sure || (need && (nzi | nzj) != 0 && !(dot < thr));
  }
  // register compare
  uint32_t mism = 0;
  if (pass) {
#pragma unroll
    for (int x = 0; x < GPL; ++x) {
      if (8u * x + q < (uint32_t)W) {
        const uint32_t av[8] = {pa[2 * x].x, pa[2 * x].y, pa[2 * x].z, pa[2 * x].w,
                                pa[2 * x + 1].x, pa[2 * x + 1].y, pa[2 * x + 1].z, pa[2 * x + 1].w};
        uint32_t bv[8] = {pb[2 * x].x, pb[2 * x].y, pb[2 * x].z, pb[2 * x].w,
                          pb[2 * x + 1].x, pb[2 * x + 1].y, pb[2 * x + 1].z, pb[2 * x + 1].w};
        column_fix(bv);
        mism += __popc(word_mismatch(av, bv));
      }
    }
  }
  mism += dpp_u32<kDppXor1>(mism);
  mism += dpp_u32<kDppXor2>(mism);        // every lane: its quad's sum
  mism += dpp_u32<kDppHalfMirror>(mism);  // plus the other quad's
  if (live && q == 0) cand[t] = pass ? (e & ~0xFFFFull) | (uint64_t)(32u * W - mism) : kDeadPair;
  ri = i;
  rj = j;
  return live && q == 0 && pass;
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b on XCD b %
// 8).  A kernel that walks a list in order gives workgroup b the contiguous
// segment xcd_segment(b) of it, so the workgroups of one XCD hold adjacent
// segments and the rows those touch stay in that XCD's L2 (gridDim.x is a
// multiple of 8).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ void xcd_segment(uint64_t total, uint64_t& s0, uint64_t& s1) {
  const uint64_t G = gridDim.x;
  const uint64_t L = (uint64_t)(blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds;
  s0 = total * L / G;
  s1 = total * (L + 1) / G;
}

// Six waves per SIMD for S <= 32: the <2, 16> instance (m = 500) needs 81
// VGPRs unconstrained (five waves) and fits 80 without spilling; the extra
// wave of gathers in flight measured 5% faster.  Wider rows would spill.
//
// A survivor (i, j) at candidate t takes the next entry of row i's list and
// of row j's: deg[i] and deg[j] count the entries, and slot[2t], slot[2t+1]
// receive the survivor's slots in row i and row j (the scatter then needs no
// atomics).  The counters are claimed in two levels.  A workgroup counts its
// survivors' entries per row in an LDS hash table (kClaimTable slots, open
// addressing); an entry takes its rank among its row's entries from an LDS
// atomic.  Every kClaimSteps steps, and at its end, the workgroup adds each
// row's count to deg with one global atomic, whose return value is the base
// of the row's ranks, and writes base + rank for its queued entries.  Device
// atomics on this part are executed at the memory side, at a limited rate
// (about 30 G/s): one atomic per entry cost the kernel 2.1 ms at 300k
// contigs.  The rows of one prefilter tile share many partners, so the table
// turns the per-entry atomics into far fewer per-row ones (63.6M into 9.3M
// there).  The atomics of a round are all issued before the first is waited
// for, since a wave waits for its loads and atomics through one counter.
constexpr uint32_t kClaimTable = 2048;               // hash slots per workgroup
constexpr uint32_t kClaimQueue = 1024;               // entries between two rounds
constexpr uint32_t kClaimSteps = kClaimQueue / 64;  // a step queues at most 64 entries
constexpr uint32_t kClaimEmpty = 0xFFFFFFFFu;

// Slot of key in the table (inserted when absent) and the entry's rank among
// the key's entries.  The table never holds more than kClaimQueue < kClaimTable
// keys, so the probe ends.
__device__ __forceinline__ uint32_t claim_insert(uint32_t* tkey, uint32_t* tval, uint32_t key,
                                                 uint32_t& rank) {
  uint32_t h = (key * 0x9E3779B1u) >> (32 - 11);  // 11 bits: kClaimTable slots
  while (true) {
    const uint32_t old = atomicCAS(&tkey[h], kClaimEmpty, key);
    if (old == kClaimEmpty || old == key) break;
    h = (h + 1u) & (kClaimTable - 1u);
  }
  rank = atomicAdd(&tval[h], 1u);
  return h;
}

template <int NC, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NC <= 2 ? 6 : 1))) void gate_compare_kernel(
    uint64_t* __restrict__ cand, uint64_t cap, const unsigned long long* __restrict__ ncand,
    const float* __restrict__ unit, const uint8_t* __restrict__ anynz,
    const AsmDesc* __restrict__ desc, const uint16_t* __restrict__ asm_of, uint32_t n_asm,
    const uint32_t* __restrict__ planes, uint32_t* __restrict__ deg, uint32_t* __restrict__ slot) {
  static_assert(kClaimTable == 2048 && kClaimTable % 256 == 0, "hash width");
  __shared__ uint32_t tkey[kClaimTable], tval[kClaimTable];
  __shared__ uint2 queue[kClaimQueue];  // (table slot | rank << 16, slot index)
  __shared__ uint32_t qn;
  const uint32_t tid = threadIdx.x;
#pragma unroll
  for (uint32_t k = 0; k < kClaimTable / 256; ++k) {
    tkey[tid + 256 * k] = kClaimEmpty;
    tval[tid + 256 * k] = 0;
  }
  if (tid == 0) qn = 0;
  __syncthreads();
  const uint64_t nc = min((uint64_t)*ncand, cap);
  uint64_t s0, s1;
  xcd_segment(nc, s0, s1);
  if (s0 >= s1) return;  // workgroup-uniform
  const uint32_t q = threadIdx.x & 7u, ln = lane_id();
  // one assembly: its gate parameters are uniform
  const bool single = n_asm == 1;
  const float thr0 = desc[0].thr;
  const uint32_t acc1_end0 = desc[0].acc1_end;
  // one round of claims: a global atomic per counter in the table, all issued
  // before the first is waited for, then base + rank for the queued entries
  auto round = [&]() {
    __syncthreads();  // entries inserted and queued
    const uint32_t n = qn;
    uint32_t base[kClaimTable / 256];
#pragma unroll
    for (uint32_t k = 0; k < kClaimTable / 256; ++k) {
      const uint32_t key = tkey[tid + 256 * k];
      base[k] = 0;
      if (key != kClaimEmpty) base[k] = atomicAdd(&deg[key], tval[tid + 256 * k]);
    }
#pragma unroll
    for (uint32_t k = 0; k < kClaimTable / 256; ++k) tval[tid + 256 * k] = base[k];
    __syncthreads();
    for (uint32_t x = tid; x < n; x += 256) {
      const uint2 en = queue[x];
      slot[en.y] = tval[en.x & 0xFFFFu] + (en.x >> 16);
    }
    __syncthreads();  // bases read before the table is cleared
#pragma unroll
    for (uint32_t k = 0; k < kClaimTable / 256; ++k) {
      tkey[tid + 256 * k] = kClaimEmpty;
      tval[tid + 256 * k] = 0;
    }
    if (tid == 0) qn = 0;
    __syncthreads();
  };
  // workgroup-uniform loop over the segment: 8 candidates per wave, 32 per
  // workgroup per step; each step's candidate is loaded one step ahead
  // (loads without branches: past the segment the last candidate is read
  // and not used)
  uint64_t tb = s0;
  uint64_t e_next = cand[min(tb + (tid >> 3), s1 - 1)];
  uint32_t steps = 0;
  for (; tb < s1; tb += 32) {
    const uint64_t t = tb + (tid >> 3);
    const uint64_t e = t < s1 ? e_next : 0;
    e_next = cand[min(t + 32, s1 - 1)];
    uint32_t ri, rj;
    const bool add = gate_compare_one<NC, W>(cand, t, e, t < s1, q, unit, anynz, desc, asm_of, single,
                                             thr0, acc1_end0, planes, ri, rj);
    uint32_t hi = 0, hj = 0, ki = 0, kj = 0;
    if (add) {
      hi = claim_insert(tkey, tval, ri, ki);
      hj = claim_insert(tkey, tval, rj, kj);
    }
    const uint64_t mask = __ballot(add);
    if (mask) {  // wave-uniform: one LDS atomic reserves the wave's queue entries
      uint32_t at = 0;
      if (ln == (uint32_t)(__ffsll((long long)mask) - 1))
        at = atomicAdd(&qn, 2u * (uint32_t)__popcll(mask));
      at = __shfl(at, __ffsll((long long)mask) - 1) + 2u * lanes_below(mask);
      if (add) {
        const uint32_t t2 = 2u * (uint32_t)t;  // slot index of row i's entry
        queue[at] = make_uint2(hi | ki << 16, t2);
        queue[at + 1] = make_uint2(hj | kj << 16, t2 + 1u);
      }
    }
    if (++steps == kClaimSteps) {
      round();
      steps = 0;
    }
  }
  round();
}

// Scatter each gated pair into both rows' candidate lists, entry =
// (cnt << 32) | partner, at the slots the exact gate claimed for it.
__global__ __launch_bounds__(256) void scatter_kernel(const uint64_t* __restrict__ surv, uint64_t cap,
                                                      const unsigned long long* __restrict__ ncand,
                                                      const uint2* __restrict__ slot,
                                                      const uint64_t* __restrict__ off,
                                                      uint64_t* __restrict__ list) {
  const uint64_t ns = min((uint64_t)*ncand, cap);
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < ns; t += (uint64_t)gridDim.x * 256) {
    const uint64_t e = surv[t];
    if (e == kDeadPair) continue;
    uint32_t i, j, c;
    unpack_pair(e, i, j, c);
    const uint2 sl = slot[t];
    const uint64_t hi = (uint64_t)c << 32;
    list[off[i] + sl.x] = hi | j;
    list[off[j] + sl.y] = hi | i;
  }
}

// ---------------------------------------------------------------------------
// Incremental builds of a served session (fused regime).  The gated pairs of
// the session's earlier rows are not evaluated again: their candidate lists,
// kept from the earlier build with local ids (CSR), are merged into this
// build's lists.  An old row's degree starts at its old list length, before
// the exact gate claims the slots of the new pairs, so those slots follow the
// old entries; the old entries then fill the head of the row's new list (ids
// made global).  Lists are sets: the selection does not depend on the order
// of a row's entries.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void merge_deg_kernel(const MergeDesc* __restrict__ md,
                                                        uint32_t* __restrict__ deg) {
  const MergeDesc d = md[blockIdx.y];
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r < d.n_old) deg[d.row0 + r] += (uint32_t)(d.loff[r + 1] - d.loff[r]);
}

// One wave per old row.
__global__ __launch_bounds__(256) void merge_lists_kernel(const MergeDesc* __restrict__ md,
                                                          const uint64_t* __restrict__ off,
                                                          uint64_t* __restrict__ list) {
  const MergeDesc d = md[blockIdx.y];
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63u;
  if (r >= d.n_old) return;
  const uint64_t s0 = d.loff[r], s1 = d.loff[r + 1];
  const uint32_t g = d.row0 + r;
  uint64_t* __restrict__ dst = list + off[g];
  // entry = count << 32 | id: adding row0 (id + row0 < 2^24) only changes the id
  for (uint64_t t = s0 + ln; t < s1; t += 64) dst[t - s0] = d.list[t] + d.row0;
}

// Rows [row0, row0 + n) of a build's lists as a local CSR: loff[r] =
// off[row0 + r] - off[row0] (r <= n), entries with ids made local.  One wave
// per row.
__global__ __launch_bounds__(256) void extract_lists_kernel(const uint64_t* __restrict__ off,
                                                            const uint64_t* __restrict__ list,
                                                            uint32_t row0, uint32_t n,
                                                            uint64_t* __restrict__ loff,
                                                            uint64_t* __restrict__ dst) {
  const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6), ln = threadIdx.x & 63u;
  if (r > n) return;
  const uint64_t base = off[row0];
  if (ln == 0) loff[r] = off[row0 + r] - base;
  if (r == n) return;
  const uint64_t s0 = off[row0 + r], s1 = off[row0 + r + 1];
  for (uint64_t t = s0 + ln; t < s1; t += 64) dst[t - base] = list[t] - row0;
}

// ---------------------------------------------------------------------------
// Per-row top-K under (sv desc, id asc) (src/rabbitbin.h:418-429).
// sv is a monotone function of the count, so the order is (level desc, id
// asc) with level = dense rank of the distinct float similarities.  A row's
// candidates are a dense matrix row (small regime, partners j != i passing
// is_nz) or a candidate list (fused regime).  Selection: level histogram ->
// threshold level L*; among candidates at L* the smallest ids are chosen by a
// two-step radix select on the id (12 + 12 bits).  The kept entries are then
// sorted by id.  Also reports the row's maximum count (small regime: over all
// partners j != i, as calibrate_sim_cutoff; fused: over the gated partners).
// ---------------------------------------------------------------------------
// The partners of one row (global row i): a row of its assembly's dense
// count matrix (regime 1; partners are all rows j != i of the assembly that
// pass is_nz) or its candidate list (regime 2).  Ids are global rows.
struct RowSource {
  const uint64_t* list;   // candidate lists
  const uint64_t* off;
  const uint16_t* mrow;   // dense: the row's counts against the assembly's rows
  const uint8_t* anynz;
  uint32_t row0, local, n;
  bool dense;
  bool small;             // regime 1: the row maximum runs over every partner
};

__device__ __forceinline__ RowSource row_source(const uint64_t* list, const uint64_t* off,
                                                const uint16_t* mat, const uint8_t* anynz,
                                                const AsmDesc& d, uint32_t i) {
  RowSource s;
  s.list = list;
  s.off = off;
  s.anynz = anynz;
  s.row0 = d.row0;
  s.local = i - d.row0;
  s.n = d.n;
  s.dense = d.dense != 0;
  s.small = d.regime == 1;
  s.mrow = s.dense ? mat + d.mat0 + (size_t)s.local * d.n : nullptr;
  return s;
}

__device__ __forceinline__ uint32_t row_len(const RowSource& s, uint32_t i) {
  return s.dense ? s.n : (uint32_t)(s.off[i + 1] - s.off[i]);
}
// Returns false when position t of row i is not a candidate.
__device__ __forceinline__ bool row_entry(const RowSource& s, uint32_t i, uint32_t t, uint32_t& id,
                                          uint32_t& c, bool& for_max) {
  if (s.dense) {
    id = s.row0 + t;
    c = s.mrow[t];
    const bool nz = !s.anynz || s.anynz[i] || s.anynz[id];
    // the small regime's calibration takes every partner (calibrate_sim_cutoff,
    // src/impl/rb_graph.cpp:1362-1384); the fused pass only evaluated ones
    for_max = t != s.local && (s.small || nz);
    return t != s.local && nz;
  }
  const uint64_t e = s.list[s.off[i] + t];
  id = (uint32_t)e;
  c = (uint32_t)(e >> 32);
  for_max = true;
  return true;
}

// Block-parallel search over an LDS histogram: the first position q of the
// scan order (ascending bins, or descending when rev) at which the inclusive
// running count reaches target.  Thread t sums a contiguous chunk of the
// scan order, a block scan gives each chunk's starting count, and the one
// thread whose chunk crosses target walks it.  Results go to *bin / *before
// (count strictly before that bin).  Requires target <= total; 256 threads.
__device__ void find_reaching(const uint32_t* h, uint32_t nb, bool rev, uint32_t target,
                              uint32_t* wave_tot, uint32_t* bin, uint32_t* before) {
  const uint32_t tid = threadIdx.x;
  const uint32_t chunk = (nb + 255u) / 256u;
  const uint32_t q0 = min(nb, tid * chunk), q1 = min(nb, q0 + chunk);
  uint32_t local = 0;
  for (uint32_t q = q0; q < q1; ++q) local += h[rev ? nb - 1 - q : q];
  uint32_t incl = local;
  const uint32_t ln = tid & 63u;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(incl, d);
    if (ln >= d) incl += t;
  }
  if (ln == 63) wave_tot[tid >> 6] = incl;
  __syncthreads();
  uint32_t excl = incl - local;
  for (uint32_t w = 0; w < (tid >> 6); ++w) excl += wave_tot[w];
  if (local != 0 && excl < target && target <= excl + local) {
    uint32_t acc = excl;
    for (uint32_t q = q0; q < q1; ++q) {
      const uint32_t b = rev ? nb - 1 - q : q;
      if (acc + h[b] >= target) {
        *bin = b;
        *before = acc;
        break;
      }
      acc += h[b];
    }
  }
  __syncthreads();
}

// Rows of a candidate list with at most kSelectStage entries: the row is
// staged once in LDS (ids, level << 16 | count); the threshold level comes
// from a level histogram; entries above it are kept, and of the entries at
// it the smallest ids fill the remaining slots (rank by id among those
// entries only).  The kept entries (at most K) are compacted and each one's
// output slot is the number of kept ids below it (ids within a row are
// distinct), counted four ids per 128-bit LDS read, so the list comes out
// sorted by id without a sorting network.  `lds` holds 4096 words: staged ids
// [0, 1024), level|count [1024, 2048), level histogram [2048, 3072), tie ids
// [3072, 4096); kept ids and counts go to sel_id / sel_cnt (>= K + 3 slots).
constexpr uint32_t kSelectStage = 1024;

__device__ __forceinline__ uint32_t count_below(const uint32_t* __restrict__ v, uint32_t nv4,
                                                uint32_t key) {
  uint32_t r = 0;
  const uint4* v4 = reinterpret_cast<const uint4*>(v);
  uint32_t q = 0;
  for (; q + 4 <= nv4; q += 4) {
    const uint4 a = v4[q], b = v4[q + 1], c = v4[q + 2], d = v4[q + 3];
    r += (a.x < key) + (a.y < key) + (a.z < key) + (a.w < key);
    r += (b.x < key) + (b.y < key) + (b.z < key) + (b.w < key);
    r += (c.x < key) + (c.y < key) + (c.z < key) + (c.w < key);
    r += (d.x < key) + (d.y < key) + (d.z < key) + (d.w < key);
  }
  for (; q < nv4; ++q) {
    const uint4 a = v4[q];
    r += (a.x < key) + (a.y < key) + (a.z < key) + (a.w < key);
  }
  return r;
}

__device__ void select_staged_row(const RowSource& src, uint32_t i, uint32_t len,
                                  const uint16_t* __restrict__ level, uint32_t nlevels, uint32_t K,
                                  uint32_t* lds, uint32_t* sel_id, uint16_t* sel_cnt,
                                  uint32_t* wave_tot, uint32_t* tmp_bin, uint32_t* tmp_below,
                                  uint32_t* rmax, uint32_t* nkept, uint32_t* ntie,
                                  uint32_t* __restrict__ topk_id, uint16_t* __restrict__ topk_cnt,
                                  uint32_t* __restrict__ topk_len,
                                  uint32_t* __restrict__ rowmax_out, uint32_t* __restrict__ fence,
                                  uint32_t FW) {
  const uint32_t tid = threadIdx.x;
  uint32_t* eid = lds;
  uint32_t* elc = lds + 1024;
  uint32_t* lhist = lds + 2048;
  uint32_t* tie = lds + 3072;
  for (uint32_t b = tid; b < 1024; b += 256) lhist[b] = 0;
  if (tid == 0) {
    *rmax = 0;
    *nkept = 0;
    *ntie = 0;
  }
  __syncthreads();
  uint32_t my_max = 0;
  for (uint32_t t = tid; t < len; t += 256) {
    const uint64_t e = src.list[src.off[i] + t];
    const uint32_t c = (uint32_t)(e >> 32), lv = level[c];
    eid[t] = (uint32_t)e;
    elc[t] = (lv << 16) | c;
    atomicAdd(&lhist[lv], 1u);
    my_max = c > my_max ? c : my_max;
  }
  if (my_max) atomicMax(rmax, my_max);
  __syncthreads();
  uint32_t Ls = 0, need_here = 0xFFFFFFFFu;  // len <= K: every entry is kept
  if (len > K) {
    find_reaching(lhist, nlevels, true, K, wave_tot, tmp_bin, tmp_below);
    Ls = *tmp_bin;
    need_here = K - *tmp_below;
  }
  const bool tie_cut = len > K && need_here < lhist[Ls];
  for (uint32_t t = tid; t < len; t += 256) {
    const uint32_t lv = elc[t] >> 16;
    if (lv > Ls || (lv == Ls && !tie_cut)) {
      const uint32_t s = atomicAdd(nkept, 1u);
      sel_id[s] = eid[t];
      sel_cnt[s] = (uint16_t)(elc[t] & 0xFFFFu);
    } else if (lv == Ls) {
      tie[atomicAdd(ntie, 1u)] = t;
    }
  }
  __syncthreads();
  if (tie_cut) {
    // the smallest need_here ids among the entries at the threshold level
    // (their counts may differ: several counts can share one level)
    const uint32_t nt = *ntie;
    uint32_t* tie_id = lhist;  // the level histogram is no longer read
    for (uint32_t u = tid; u < ((nt + 3) & ~3u); u += 256) tie_id[u] = u < nt ? eid[tie[u]] : 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t u = tid; u < nt; u += 256) {
      const uint32_t me = tie_id[u];
      if (count_below(tie_id, (nt + 3) / 4, me) < need_here) {
        const uint32_t s = atomicAdd(nkept, 1u);
        sel_id[s] = me;
        sel_cnt[s] = (uint16_t)(elc[tie[u]] & 0xFFFFu);
      }
    }
    __syncthreads();
  }
  const uint32_t ns = *nkept;  // = min(len, K)
  for (uint32_t u = ns + tid; u < ((ns + 3) & ~3u); u += 256) sel_id[u] = 0xFFFFFFFFu;
  __syncthreads();
  for (uint32_t t = tid; t < ns; t += 256) {
    const uint32_t me = sel_id[t];
    const uint32_t pos = count_below(sel_id, (ns + 3) / 4, me);
    topk_id[(size_t)i * K + pos] = me;
    topk_cnt[(size_t)i * K + pos] = sel_cnt[t];
    if ((pos & 15u) == 0) fence[(size_t)i * FW + pos / 16] = me;
  }
  for (uint32_t q = (ns + 15) / 16 + tid; q < FW; q += 256) fence[(size_t)i * FW + q] = 0xFFFFFFFFu;
  if (tid == 0) {
    topk_len[i] = ns;
    rowmax_out[i] = *rmax;
  }
}

// Block-wide exclusive prefix sum of one value per thread (256 threads);
// `total` receives the block's sum.  wtot: 4 words of LDS.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* wtot, uint32_t& total) {
  const uint32_t ln = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(incl, d);
    if (ln >= d) incl += t;
  }
  if (ln == 63) wtot[w] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t x = wtot[k];
    before += k < w ? x : 0u;
    tot += x;
  }
  __syncthreads();  // wtot is reused by the next call
  total = tot;
  return before + incl - v;
}

// Dense rows (the small regime, and the fused regime without the abundance
// gate): the row's counts against every row of its assembly, in id order;
// the diagonal and partners failing is_nz are skipped.  Pass 1 reads the row
// eight counts per 128-bit load and builds the count histogram (counts at
// level 0, the similarity clamped to 0, are counted in registers) and the
// row maximum; the threshold level L* and the number `need` of entries to
// take at L* follow from the histogram folded into levels.  Pass 2 keeps, in
// id order, every entry above L* and the first `need` entries at L*: an
// ordered compaction with block-wide prefix counts per 2048-entry step, so
// the list comes out sorted by id with no sort.  The row's counts are read
// through 16-byte aligned chunks, so the matrix buffer keeps 16 readable
// bytes past its last row.  `lds` holds 4096 words.
__device__ void select_dense_row(const RowSource& src, uint32_t i, const uint16_t* __restrict__ level_g,
                                 uint32_t m, uint32_t K, uint32_t* lds, uint32_t* wtot, uint32_t* tmp_bin,
                                 uint32_t* tmp_below, uint32_t* rmax, uint32_t* ztot,
                                 uint32_t* __restrict__ topk_id, uint16_t* __restrict__ topk_cnt,
                                 uint32_t* __restrict__ topk_len, uint32_t* __restrict__ rowmax_out,
                                 uint32_t* __restrict__ fence, uint32_t FW) {
  const uint32_t tid = threadIdx.x, wv = tid >> 6;
  uint32_t* lvl = lds;                       // level of each count, [0, m]
  uint32_t* whist = lds + 1024 + 520 * wv;   // this wave's count histogram
  uint32_t* lhist = lds + 3104;              // level histogram, [0, m]
  const uint32_t nb = m + 1;
  for (uint32_t c = tid; c < nb; c += 256) {
    lvl[c] = level_g[c];
    lhist[c] = 0;
  }
  for (uint32_t c = tid; c < 4 * 520; c += 256) lds[1024 + c] = 0;
  if (tid == 0) {
    *rmax = 0;
    *ztot = 0;
  }
  __syncthreads();
  const uint32_t n = src.n, local = src.local, row0 = src.row0;
  const uint16_t* __restrict__ row = src.mrow;
  const uint32_t head = (uint32_t)(((uintptr_t)row & 15u) >> 1);  // counts before the row in its first chunk
  const uint4* __restrict__ chunks = reinterpret_cast<const uint4*>(row - head);
  const uint32_t nch = (head + n + 7u) >> 3;
  const bool nz_i = !src.anynz || src.anynz[i];  // then every partner passes is_nz
  // one chunk's entries: valid (a partner of the row) and is_nz
  auto entry = [&](uint32_t ch, uint32_t e, uint32_t word, uint32_t& c, bool& ok, bool& fm) {
    const int64_t t = (int64_t)ch * 8 + e - head;
    c = (word >> (16 * (e & 1u))) & 0xFFFFu;
    const bool in = t >= 0 && t < (int64_t)n && (uint32_t)t != local;
    const bool nz = nz_i || (in && src.anynz[row0 + (uint32_t)t]);
    ok = in && nz;
    fm = in && (src.small || nz);
  };
  uint32_t my_max = 0, zeros = 0;
  for (uint32_t ch = tid; ch < nch; ch += 256) {
    const uint4 v = chunks[ch];
    const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) {
      uint32_t c;
      bool ok, fm;
      entry(ch, e, wd[e >> 1], c, ok, fm);
      if (fm && c > my_max) my_max = c;
      if (ok) {
        if (lvl[c] == 0)
          ++zeros;
        else
          atomicAdd(&whist[c], 1u);
      }
    }
  }
  if (my_max) atomicMax(rmax, my_max);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) zeros += __shfl_xor(zeros, o);
  if ((tid & 63u) == 0 && zeros) atomicAdd(ztot, zeros);
  __syncthreads();
  for (uint32_t c = tid; c < nb; c += 256) {
    const uint32_t k = lds[1024 + c] + lds[1544 + c] + lds[2064 + c] + lds[2584 + c];
    if (k) atomicAdd(&lhist[lvl[c]], k);
  }
  __syncthreads();
  if (tid == 0) lhist[0] += *ztot;
  __syncthreads();
  uint32_t part = 0;
  for (uint32_t b = tid; b < nb; b += 256) part += lhist[b];
  uint32_t total;
  (void)block_scan_excl(part, wtot, total);
  const bool take_all = total <= K;
  uint32_t Ls = 0, need = 0xFFFFFFFFu;
  if (!take_all) {
    find_reaching(lhist, nb, true, K, wtot, tmp_bin, tmp_below);
    Ls = *tmp_bin;
    need = K - *tmp_below;
  }
  const uint32_t ns = take_all ? total : K;
  // pass 2: ordered compaction
  uint32_t out = 0, at_ls = 0;  // kept so far, entries at L* seen so far (block-uniform)
  for (uint32_t step = 0; step * 256 < nch && out < ns; ++step) {
    const uint32_t ch = step * 256 + tid;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (ch < nch) v = chunks[ch];
    const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
    uint32_t gt = 0, eq = 0;  // bit e: entry e above L* / at L*
#pragma unroll
    for (uint32_t e = 0; e < 8; ++e) {
      uint32_t c;
      bool ok, fm;
      entry(ch, e, wd[e >> 1], c, ok, fm);
      ok = ok && ch < nch;
      const uint32_t lv = ok ? lvl[c] : 0u;  // counts outside the row are not read as table indices
      if (ok && (take_all || lv > Ls)) gt |= 1u << e;
      if (ok && !take_all && lv == Ls) eq |= 1u << e;
    }
    uint32_t eq_total;
    const uint32_t eq_before = at_ls + block_scan_excl(__popc(eq), wtot, eq_total);
    // entries at L* whose rank among them is below need are kept
    uint32_t keep = gt;
    for (uint32_t b = eq, r = eq_before; b; b &= b - 1u, ++r)
      if (r < need) keep |= 1u << (__ffs(b) - 1u);
    uint32_t keep_total;
    uint32_t pos = out + block_scan_excl(__popc(keep), wtot, keep_total);
    for (uint32_t b = keep; b; b &= b - 1u, ++pos) {
      const uint32_t e = __ffs(b) - 1u;
      const uint32_t t = ch * 8 + e - head;
      const uint32_t c = (wd[e >> 1] >> (16 * (e & 1u))) & 0xFFFFu;
      topk_id[(size_t)i * K + pos] = row0 + t;
      topk_cnt[(size_t)i * K + pos] = (uint16_t)c;
      if ((pos & 15u) == 0) fence[(size_t)i * FW + pos / 16] = row0 + t;
    }
    out += keep_total;
    at_ls += eq_total;
  }
  for (uint32_t q = (ns + 15) / 16 + tid; q < FW; q += 256) fence[(size_t)i * FW + q] = 0xFFFFFFFFu;
  if (tid == 0) {
    topk_len[i] = ns;
    rowmax_out[i] = *rmax;
  }
}

// One row with the whole workgroup: dense rows, and candidate lists longer
// than select_wave_kernel takes.
__device__ __forceinline__ void select_block_row(
    uint32_t i, const uint64_t* __restrict__ list, const uint64_t* __restrict__ off,
    const uint16_t* __restrict__ mat, const uint8_t* __restrict__ anynz,
    const AsmDesc* __restrict__ desc, const uint16_t* __restrict__ asm_of,
    const uint16_t* __restrict__ level_all, uint32_t m, uint32_t K,
    uint32_t* __restrict__ topk_id, uint16_t* __restrict__ topk_cnt,
    uint32_t* __restrict__ topk_len, uint32_t* __restrict__ rowmax_out,
    uint32_t* __restrict__ fence, uint32_t FW) {
  __shared__ __align__(16) uint32_t hist[4096];
  __shared__ __align__(16) uint32_t sel_id[1024];
  __shared__ uint16_t sel_cnt[1024];
  __shared__ uint32_t nsel, tot, tmp_bin, tmp_below, rmax;
  __shared__ uint32_t wave_tot[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t a = asm_of[i];
  const RowSource src = row_source(list, off, mat, anynz, desc[a], i);
  // the assembly's level table; bins above its distinct similarities stay empty
  const uint16_t* __restrict__ level = level_all + (size_t)a * (m + 1);
  const uint32_t nlevels = m + 1;
  const uint32_t len = row_len(src, i);

  if (src.dense && nlevels <= 520) {
    select_dense_row(src, i, level, m, K, hist, wave_tot, &tmp_bin, &tmp_below, &rmax, &tot, topk_id,
                     topk_cnt, topk_len, rowmax_out, fence, FW);
    return;
  }
  if (!src.dense && len <= kSelectStage && nlevels <= 1024) {
    select_staged_row(src, i, len, level, nlevels, K, hist, sel_id, sel_cnt, wave_tot, &tmp_bin,
                      &tmp_below, &rmax, &nsel, &tot, topk_id, topk_cnt, topk_len, rowmax_out,
                      fence, FW);
    return;
  }
  for (uint32_t b = tid; b < 4096; b += 256) hist[b] = 0;
  if (tid == 0) {
    nsel = 0;
    rmax = 0;
    tot = 0;
  }
  __syncthreads();
  // Level histogram.  Level 0 (similarity clamped to 0) holds about half of a
  // long row: it is counted in registers; the other levels go to one LDS
  // histogram per wave (1024 bins each), summed after.
  uint32_t my_max = 0, zeros = 0;
  uint32_t* __restrict__ whist = hist + 1024 * (tid >> 6);
  for (uint32_t t = tid; t < len; t += 256) {
    uint32_t id, c;
    bool fm;
    const bool ok = row_entry(src, i, t, id, c, fm);
    if (fm && c > my_max) my_max = c;
    if (ok) {
      const uint32_t lv = level[c];
      if (lv == 0)
        ++zeros;
      else
        atomicAdd(&whist[lv], 1u);
    }
  }
  if (my_max) atomicMax(&rmax, my_max);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) zeros += __shfl_xor(zeros, o);
  if ((tid & 63u) == 0 && zeros) atomicAdd(&hist[3072], zeros);  // every wave's bin 0 is empty
  __syncthreads();
  for (uint32_t b = tid; b < nlevels; b += 256)  // bins < 1024: each read and written by one thread
    hist[b] = hist[b] + hist[1024 + b] + hist[2048 + b] + hist[3072 + b];
  // threshold level L*: the largest level whose suffix count reaches K
  __syncthreads();
  {
    uint32_t part = 0;
    for (uint32_t b = tid; b < nlevels; b += 256) part += hist[b];
    if (part) atomicAdd(&tot, part);
  }
  __syncthreads();
  const bool take_all = tot <= K;
  uint32_t Ls = 0xFFFFFFFFu, need_here = 0;
  if (!take_all) {
    find_reaching(hist, nlevels, true, K, wave_tot, &tmp_bin, &tmp_below);
    Ls = tmp_bin;
    need_here = K - tmp_below;
  }
  uint32_t idthr = 0xFFFFFFFFu;  // keep ids <= idthr at level L*
  if (!take_all && need_here < hist[Ls]) {
    // radix select on the id among level == L*: high 12 bits, then low 12
    __syncthreads();
    for (uint32_t b = tid; b < 4096; b += 256) hist[b] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < len; t += 256) {
      uint32_t id, c;
      bool fm;
      if (row_entry(src, i, t, id, c, fm) && level[c] == Ls) atomicAdd(&hist[(id >> 12) & 4095u], 1u);
    }
    __syncthreads();
    find_reaching(hist, 4096, false, need_here, wave_tot, &tmp_bin, &tmp_below);
    const uint32_t hb = tmp_bin;
    const uint32_t need2 = need_here - tmp_below;
    __syncthreads();
    for (uint32_t b = tid; b < 4096; b += 256) hist[b] = 0;
    __syncthreads();
    for (uint32_t t = tid; t < len; t += 256) {
      uint32_t id, c;
      bool fm;
      if (row_entry(src, i, t, id, c, fm) && level[c] == Ls && ((id >> 12) & 4095u) == hb)
        atomicAdd(&hist[id & 4095u], 1u);
    }
    __syncthreads();
    find_reaching(hist, 4096, false, need2, wave_tot, &tmp_bin, &tmp_below);
    idthr = (hb << 12) | tmp_bin;
  }
  // collect
  for (uint32_t t = tid; t < len; t += 256) {
    uint32_t id, c;
    bool fm;
    if (!row_entry(src, i, t, id, c, fm)) continue;
    const uint32_t lv = level[c];
    bool keep;
    if (take_all)
      keep = true;
    else
      keep = lv > Ls || (lv == Ls && id <= idthr);
    if (keep) {
      const uint32_t s = atomicAdd(&nsel, 1u);
      if (s < 1024) {
        sel_id[s] = id;
        sel_cnt[s] = (uint16_t)c;
      }
    }
  }
  __syncthreads();
  const uint32_t ns = min(nsel, K);
  // bitonic sort by id over the next power of two >= ns
  uint32_t P = 1;
  while (P < ns) P <<= 1;
  for (uint32_t t = ns + tid; t < P; t += 256) {
    sel_id[t] = 0xFFFFFFFFu;
    sel_cnt[t] = 0;
  }
  __syncthreads();
  for (uint32_t k2 = 2; k2 <= P; k2 <<= 1) {
    for (uint32_t jj = k2 >> 1; jj > 0; jj >>= 1) {
      for (uint32_t t = tid; t < P; t += 256) {
        const uint32_t u = t ^ jj;
        if (u > t) {
          const bool asc = (t & k2) == 0;
          const uint32_t a = sel_id[t], bb = sel_id[u];
          if ((a > bb) == asc) {
            sel_id[t] = bb;
            sel_id[u] = a;
            const uint16_t ca = sel_cnt[t];
            sel_cnt[t] = sel_cnt[u];
            sel_cnt[u] = ca;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t t = tid; t < ns; t += 256) {
    topk_id[(size_t)i * K + t] = sel_id[t];
    topk_cnt[(size_t)i * K + t] = sel_cnt[t];
  }
  for (uint32_t q = tid; q < FW; q += 256)
    fence[(size_t)i * FW + q] = 16 * q < ns ? sel_id[16 * q] : 0xFFFFFFFFu;
  if (tid == 0) {
    topk_len[i] = ns;
    rowmax_out[i] = rmax;
  }
}

// One workgroup per row: workgroup b takes queue[b], a dense row the plan
// listed (one row per workgroup and no loop, so the row index stays in
// scalar registers and the kernel keeps the row function's occupancy).
__global__ __launch_bounds__(256) void select_topk_kernel(
    const uint32_t* __restrict__ queue, const uint64_t* __restrict__ list,
    const uint64_t* __restrict__ off, const uint16_t* __restrict__ mat,
    const uint8_t* __restrict__ anynz, const AsmDesc* __restrict__ desc,
    const uint16_t* __restrict__ asm_of, const uint16_t* __restrict__ level_all, uint32_t m,
    uint32_t K, uint32_t* __restrict__ topk_id, uint16_t* __restrict__ topk_cnt,
    uint32_t* __restrict__ topk_len, uint32_t* __restrict__ rowmax_out, uint32_t* __restrict__ fence,
    uint32_t FW) {
  select_block_row(queue[blockIdx.x], list, off, mat, anynz, desc, asm_of, level_all, m, K, topk_id, topk_cnt,
                   topk_len, rowmax_out, fence, FW);
}

// The long candidate lists select_wave_kernel queued (queue[nstatic + q],
// q < qcount[0]), taken one at a time from the counter qcount[1].
__global__ __launch_bounds__(256) void select_long_kernel(
    const uint32_t* __restrict__ queue, uint32_t nstatic, uint32_t* __restrict__ qcount,
    const uint64_t* __restrict__ list, const uint64_t* __restrict__ off,
    const uint16_t* __restrict__ mat, const uint8_t* __restrict__ anynz,
    const AsmDesc* __restrict__ desc, const uint16_t* __restrict__ asm_of,
    const uint16_t* __restrict__ level_all, uint32_t m, uint32_t K,
    uint32_t* __restrict__ topk_id, uint16_t* __restrict__ topk_cnt,
    uint32_t* __restrict__ topk_len, uint32_t* __restrict__ rowmax_out,
    uint32_t* __restrict__ fence, uint32_t FW) {
  __shared__ uint32_t next;
  const uint32_t nq = qcount[0];
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(&qcount[1], 1u);
    __syncthreads();
    const uint32_t q = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();  // read before thread 0 takes the next row
    if (q >= nq) break;  // workgroup-uniform
    select_block_row(queue[nstatic + q], list, off, mat, anynz, desc, asm_of, level_all, m, K, topk_id,
                     topk_cnt, topk_len, rowmax_out, fence, FW);
    __syncthreads();  // the row's LDS is read before the next row writes it
  }
}

// Candidate-list rows of at most kWaveSelect entries, one wave per row (four
// rows per workgroup) and no workgroup barrier: the same selection as
// select_staged_row.  Lane l holds entries l, l + 64, ... (id, level << 16 |
// count) in registers.  When the row has more than K entries, a level
// histogram in the wave's LDS gives the threshold level L* (the largest
// level whose suffix count reaches K): lane l sums bins 16l..16l+15, a
// suffix scan over the lanes finds the lane holding L*, and that lane walks
// its bins.  Entries above L* are kept; of those at L*, all when they fit,
// else the need smallest ids (rank among the tied ids).  The kept entries
// are compacted with ballots, and each one's output slot is the number of
// kept ids below it, so the list comes out sorted by id.  Longer lists are
// appended to `queue` (after its nstatic dense rows) for select_topk_kernel;
// dense rows are left to it.
constexpr uint32_t kWaveSelect = 512;
constexpr uint32_t kWsPer = kWaveSelect / 64;  // entries per lane

__global__ __launch_bounds__(256) void select_wave_kernel(
    const uint64_t* __restrict__ list, const uint64_t* __restrict__ off,
    const AsmDesc* __restrict__ desc, const uint16_t* __restrict__ asm_of,
    const uint16_t* __restrict__ level_all, uint32_t m, uint32_t nrows, uint32_t K,
    uint32_t* __restrict__ topk_id, uint16_t* __restrict__ topk_cnt,
    uint32_t* __restrict__ topk_len, uint32_t* __restrict__ rowmax_out,
    uint32_t* __restrict__ fence, uint32_t FW, uint32_t* __restrict__ queue, uint32_t nstatic,
    uint32_t* __restrict__ qcount) {
  // per wave: the level histogram (1024 bins), then kept ids [0, 512) and
  // tied ids [512, 1024); kept counts
  __shared__ __align__(16) uint32_t wlds[4][1024];
  __shared__ uint16_t wcnt[4][kWaveSelect];
  const uint32_t wv = threadIdx.x >> 6, ln = lane_id();
  const uint32_t i = blockIdx.x * 4 + wv;
  if (i >= nrows) return;  // wave-uniform
  const uint32_t a = asm_of[i];
  const uint64_t o0 = off[i];
  const uint32_t len = (uint32_t)(off[i + 1] - o0);
  if (desc[a].dense) return;  // listed by the plan
  if (len > kWaveSelect) {
    if (ln == 0) queue[nstatic + atomicAdd(qcount, 1u)] = i;
    return;
  }
  const uint16_t* __restrict__ level = level_all + (size_t)a * (m + 1);
  const uint32_t nlevels = m + 1;
  uint32_t* __restrict__ h = wlds[wv];
  uint32_t id[kWsPer], cl[kWsPer];
  uint32_t my_max = 0;
#pragma unroll
  for (uint32_t k = 0; k < kWsPer; ++k) {
    const uint32_t t = ln + 64 * k;
    id[k] = 0xFFFFFFFFu;
    cl[k] = 0;
    if (t < len) {
      const uint64_t e = list[o0 + t];
      const uint32_t c = (uint32_t)(e >> 32);
      id[k] = (uint32_t)e;
      cl[k] = (uint32_t)level[c] << 16 | c;
      my_max = max(my_max, c);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) my_max = max(my_max, (uint32_t)__shfl_xor((int)my_max, o));
  uint32_t Ls = 0, need = 0xFFFFFFFFu;
  bool tie_cut = false;
  if (len > K) {
    for (uint32_t b = ln; b < 1024; b += 64) h[b] = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t k = 0; k < kWsPer; ++k)
      if (ln + 64 * k < len) atomicAdd(&h[cl[k] >> 16], 1u);
    __builtin_amdgcn_wave_barrier();
    uint32_t part = 0;
#pragma unroll
    for (uint32_t q = 0; q < 16; ++q) part += h[16 * ln + q];  // bins >= nlevels are zero
    // entries in the bins of the lanes above this one
    uint32_t incl = part;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_down((int)incl, d);
      if (ln + d < 64) incl += t;
    }
    const uint32_t above = incl - part;
    const bool mine = above < K && K <= incl;  // exactly one lane
    uint32_t lb = 0, lbelow = 0;
    if (mine) {
      uint32_t acc = above;
      for (int q = 15; q >= 0; --q) {
        const uint32_t b = 16 * ln + (uint32_t)q;
        if (acc + h[b] >= K) {
          lb = b;
          lbelow = acc;
          break;
        }
        acc += h[b];
      }
    }
    const uint64_t owner = __ballot(mine);  // one lane (the row has more than K entries)
    const int src = owner ? __ffsll((long long)owner) - 1 : 0;
    Ls = (uint32_t)__shfl((int)lb, src) & 1023u;
    need = K - (uint32_t)__shfl((int)lbelow, src);
    tie_cut = need < h[Ls];
    (void)nlevels;
    __builtin_amdgcn_wave_barrier();  // the histogram is read before it is reused
  }
  uint32_t* __restrict__ kid = h;          // kept ids
  uint32_t* __restrict__ tid_ = h + 512;   // tied ids (tie_cut only)
  uint16_t* __restrict__ kc = wcnt[wv];
  uint32_t nk = 0, nt = 0;
#pragma unroll
  for (uint32_t k = 0; k < kWsPer; ++k) {
    const bool v = ln + 64 * k < len;
    const uint32_t lv = cl[k] >> 16;
    const bool keep = v && (len <= K || lv > Ls || (lv == Ls && !tie_cut));
    const bool tie = v && len > K && tie_cut && lv == Ls;
    const uint64_t km = __ballot(keep), tm = __ballot(tie);
    if (keep) {
      const uint32_t p = nk + lanes_below(km);
      kid[p] = id[k];
      kc[p] = (uint16_t)(cl[k] & 0xFFFFu);
    }
    if (tie) tid_[nt + lanes_below(tm)] = id[k];
    nk += (uint32_t)__popcll(km);
    nt += (uint32_t)__popcll(tm);
  }
  if (tie_cut) {
    for (uint32_t u = nt + ln; u < ((nt + 3) & ~3u); u += 64) tid_[u] = 0xFFFFFFFFu;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t k = 0; k < kWsPer; ++k) {
      const bool tie = ln + 64 * k < len && (cl[k] >> 16) == Ls;
      const bool take = tie && count_below(tid_, (nt + 3) / 4, id[k]) < need;
      const uint64_t km = __ballot(take);
      if (take) {
        const uint32_t p = nk + lanes_below(km);
        kid[p] = id[k];
        kc[p] = (uint16_t)(cl[k] & 0xFFFFu);
      }
      nk += (uint32_t)__popcll(km);
    }
  }
  const uint32_t ns = nk;  // = min(len, K)
  for (uint32_t u = ns + ln; u < ((ns + 3) & ~3u); u += 64) kid[u] = 0xFFFFFFFFu;
  __builtin_amdgcn_wave_barrier();
  for (uint32_t t = ln; t < ns; t += 64) {
    const uint32_t me = kid[t];
    const uint32_t pos = count_below(kid, (ns + 3) / 4, me);
    topk_id[(size_t)i * K + pos] = me;
    topk_cnt[(size_t)i * K + pos] = kc[t];
    if ((pos & 15u) == 0) fence[(size_t)i * FW + pos / 16] = me;
  }
  for (uint32_t q = (ns + 15) / 16 + ln; q < FW; q += 64) fence[(size_t)i * FW + q] = 0xFFFFFFFFu;
  if (ln == 0) {
    topk_len[i] = ns;
    rowmax_out[i] = my_max;
  }
}

// Edge emission.  Small regime: (i, j) for i < j, j in top(i), count >= cmin.
// Fused regime additionally requires i in top(j) (mutual, default on,
// src/impl/rb_graph.cpp:1283-1348).  Top-K lists are sorted by id.
// Pass 0 (emit_mark_kernel): one wave per row; lane l takes list slots
// l, l+64, l+128, l+192 together.  The membership test of i in top(j) reads
// j's fence row (the ids at list positions 0, 16, 32, ..., written by the
// selection) to find the 16-entry segment that can hold i, then that
// segment: two 64-byte reads per test.  The keep decisions are stored as one 64-bit mask per slot group
// and the per-row edge count.  Pass 1 (emit_write_kernel) writes the kept
// entries at the row's offset, so the edges come out sorted by (i, j).
__global__ __launch_bounds__(256) void emit_mark_kernel(
    const uint32_t* __restrict__ topk_id, const uint16_t* __restrict__ topk_cnt,
    const uint32_t* __restrict__ topk_len, const uint32_t* __restrict__ fence, uint32_t FW,
    uint32_t n, uint32_t K, const BuildScalars* __restrict__ bs, const AsmDesc* __restrict__ desc,
    const uint16_t* __restrict__ asm_of, const uint32_t* __restrict__ order, uint32_t KW,
    uint64_t* __restrict__ keep, uint32_t* __restrict__ row_edges) {
  const bool aligned16 = (K & 3u) == 0;
  const uint32_t pos = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t ln = threadIdx.x & 63u;
  if (pos >= n) return;
  // rows in pair-pass order: rows of one abundance cluster, whose partners
  // overlap, are tested together
  const uint32_t i = order ? order[pos] : pos;
  const uint32_t a = asm_of[i];
  const uint32_t cmin = bs[a].cmin;
  const bool mutual = desc[a].regime == 2;
  const uint32_t len = topk_len[i];
  uint32_t total = 0;
  for (uint32_t s0 = 0; s0 * 64 < len; s0 += 4) {
    uint32_t jv[4], L[4];
    bool k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t t = (s0 + u) * 64 + ln;
      k[u] = false;
      jv[u] = 0;
      if (t < len) {
        jv[u] = topk_id[(size_t)i * K + t];
        k[u] = jv[u] > i && topk_cnt[(size_t)i * K + t] >= cmin;
      }
    }
    if (mutual) {
      // fence row of j: ids at list positions 0, 16, 32, ... (0xFFFFFFFF past
      // the end); the number of fences <= i locates i's 16-entry segment
      uint32_t qn[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        L[u] = k[u] ? topk_len[jv[u]] : 0u;
        qn[u] = 0;
      }
      for (uint32_t f = 0; f < FW; f += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (k[u]) {
            const uint4 fv = *reinterpret_cast<const uint4*>(fence + (size_t)jv[u] * FW + f);
            qn[u] += (fv.x <= i) + (fv.y <= i) + (fv.z <= i) + (fv.w <= i);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (k[u]) {
          bool hit = false;
          if (qn[u] > 0) {
            const uint32_t p0 = 16 * (qn[u] - 1);
            const uint32_t* seg = topk_id + (size_t)jv[u] * K + p0;
#pragma unroll
            for (int x = 0; x < 16; x += 4) {
              const uint4 v = aligned16 ? *reinterpret_cast<const uint4*>(seg + x)
                                        : make_uint4(seg[x], seg[x + 1], seg[x + 2], seg[x + 3]);
              hit |= (v.x == i && p0 + x < L[u]) | (v.y == i && p0 + x + 1 < L[u]) |
                     (v.z == i && p0 + x + 2 < L[u]) | (v.w == i && p0 + x + 3 < L[u]);
            }
          }
          k[u] = hit;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if ((s0 + u) * 64 < len) {
        const uint64_t m = __ballot(k[u]);
        if (ln == 0) keep[(size_t)i * KW + s0 + u] = m;
        total += (uint32_t)__popcll(m);
      }
    }
  }
  if (ln == 0) row_edges[i] = total;
}

// It also writes each edge's composition similarity sComp = svf[count] of the
// row's assembly.
__global__ __launch_bounds__(256) void emit_write_kernel(
    const uint32_t* __restrict__ topk_id, const uint16_t* __restrict__ topk_cnt,
    const uint32_t* __restrict__ topk_len, uint32_t n, uint32_t K, uint32_t KW,
    const uint64_t* __restrict__ keep, const uint64_t* __restrict__ row_off,
    const uint16_t* __restrict__ asm_of, uint32_t m, const float* __restrict__ svf_all,
    uint32_t* __restrict__ efrom, uint32_t* __restrict__ eto, uint16_t* __restrict__ ecnt,
    float* __restrict__ escomp) {
  const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t ln = threadIdx.x & 63u;
  if (i >= n) return;
  const uint32_t len = topk_len[i];
  const uint64_t off = row_off[i];
  const float* __restrict__ svf = svf_all + (size_t)asm_of[i] * (m + 1);
  uint64_t written = 0;
  for (uint32_t s = 0; s * 64 < len; ++s) {
    const uint64_t m = keep[(size_t)i * KW + s];
    if ((m >> ln) & 1ull) {
      const uint32_t t = s * 64 + ln;
      const uint64_t pos = off + written + lanes_below(m);
      const uint16_t c = topk_cnt[(size_t)i * K + t];
      efrom[pos] = i;
      eto[pos] = topk_id[(size_t)i * K + t];
      ecnt[pos] = c;
      escomp[pos] = svf[c];
    }
    written += (uint32_t)__popcll(m);
  }
}

// Stage (d): per-edge weight, the reference's depth_edge_term with the
// default fuse = min (src/rabbitbin.cpp:930-970) and the combination and
// cutoff of src/rabbitbin.cpp:5500-5507.
//
// The correlation term is summed as depth_corr_fast sums it on the
// reference's AVX-512 build (src/rabbitbin.cpp:705-746): the whole 16-float
// chunks by per-lane FP32 FMAs (chunk pairs of a 32-wide step go to two
// accumulators), one reduce in GCC's _mm512_reduce_add_ps order (lanes l+8,
// then l+4, then (0,2),(1,3), then the last pair), then the remaining S % 16
// products in FP64.  Below 16 samples that is the plain FP64 sum.  The same
// sum in FP64 throughout is the FP64 weight the near-cut list reports.

// corr in the reference order over two rows given by pointers (any S).
__device__ __forceinline__ double corr_ref_ptr(const float* __restrict__ a, const float* __restrict__ b,
                                               uint32_t S) {
  float acc0[16], acc1[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) acc0[l] = acc1[l] = 0.f;
  uint32_t k = 0;
  for (; k + 32 <= S; k += 32) {
#pragma unroll
    for (int l = 0; l < 16; ++l) {
      acc0[l] = __builtin_fmaf(a[k + l], b[k + l], acc0[l]);
      acc1[l] = __builtin_fmaf(a[k + 16 + l], b[k + 16 + l], acc1[l]);
    }
  }
  for (; k + 16 <= S; k += 16) {
#pragma unroll
    for (int l = 0; l < 16; ++l) acc0[l] = __builtin_fmaf(a[k + l], b[k + l], acc0[l]);
  }
  double s = 0.0;
  if (k > 0) {
    float t8[8], t4[4];
#pragma unroll
    for (int l = 0; l < 8; ++l) t8[l] = (acc0[l + 8] + acc1[l + 8]) + (acc0[l] + acc1[l]);
#pragma unroll
    for (int l = 0; l < 4; ++l) t4[l] = t8[l + 4] + t8[l];
    s = (double)((t4[0] + t4[2]) + (t4[1] + t4[3]));
  }
  for (; k < S; ++k) s += (double)a[k] * (double)b[k];
  return s;
}

// The same over a row held in registers (N = 16 NC floats) and a staged row;
// returns the reference-order sum, and the FP64 sum in c64.
template <int N>
__device__ __forceinline__ double corr_ref_regs(const float (&a)[N], const float* __restrict__ b, uint32_t S,
                                                double& c64) {
  float acc0[16], acc1[16];
#pragma unroll
  for (int l = 0; l < 16; ++l) acc0[l] = acc1[l] = 0.f;
  const uint32_t nfull = S >> 4, n32 = 2u * (S >> 5);
#pragma unroll
  for (int c = 0; c < N / 16; ++c) {
    if ((uint32_t)c < nfull) {
      if ((c & 1) && (uint32_t)c < n32) {
#pragma unroll
        for (int l = 0; l < 16; ++l) acc1[l] = __builtin_fmaf(a[16 * c + l], b[16 * c + l], acc1[l]);
      } else {
#pragma unroll
        for (int l = 0; l < 16; ++l) acc0[l] = __builtin_fmaf(a[16 * c + l], b[16 * c + l], acc0[l]);
      }
    }
  }
  double s = 0.0;
  if (nfull > 0) {
    float t8[8], t4[4];
#pragma unroll
    for (int l = 0; l < 8; ++l) t8[l] = (acc0[l + 8] + acc1[l + 8]) + (acc0[l] + acc1[l]);
#pragma unroll
    for (int l = 0; l < 4; ++l) t4[l] = t8[l + 4] + t8[l];
    s = (double)((t4[0] + t4[2]) + (t4[1] + t4[3]));
  }
  double f = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if ((uint32_t)k < S) {
      const double p = (double)a[k] * (double)b[k];  // exact
      f += p;
      if ((uint32_t)k >= 16u * nfull) s += p;
    }
  }
  c64 = f;
  return s;
}

__device__ __forceinline__ bool finite64(double x) {
  return x == x && x != __builtin_huge_val() && x != -__builtin_huge_val();
}

// Weighted Jaccard of column-normalised raw depths (src/rabbitbin.cpp:685-698),
// clamped to [0, 1] as depth_edge_term clamps it.
__device__ __forceinline__ double wjac_clamped(double mn, double mx) {
  double wj = mx <= 0.0 ? 0.0 : mn / mx;
  if (!(wj == wj)) wj = 0.0;
  return wj < 0.0 ? 0.0 : (wj > 1.0 ? 1.0 : wj);
}

// Edge weight from the correlation and the weighted Jaccard; `pre` is the
// weight before the min-edge cut (0 when the correlation cuts the edge).
__device__ __forceinline__ double edge_weight_of(double corr, double wj, double sc, double w_comp,
                                                 double min_edge, double neg_thr, double& pre) {
  pre = 0.0;
  if (!finite64(corr) || (neg_thr > -0.99 && corr < neg_thr)) return 0.0;
  const double cp = corr < 0.0 ? 0.0 : corr;
  const double d = cp < wj ? cp : wj;
  double w = w_comp * sc + (1.0 - w_comp) * d;
  pre = w;
  if (!(w == w) || w < min_edge) w = 0.0;
  return w;
}

// Whether the FP64 weight of an edge lies within tol of a cut: its
// correlation within tol of the negative cut, or its weight before the
// min-edge cut within tol of that cut.
__device__ __forceinline__ bool near_cut(double corr64, double pre64, double min_edge, double neg_thr,
                                         double tol) {
  const bool ncorr = neg_thr > -0.99 && fabs(corr64 - neg_thr) <= tol;
  const bool nw = fabs(pre64 - min_edge) <= tol;
  return ncorr || nw;
}

// rbg_edge_weights: one edge per thread, rows unpadded (any S).  score is
// the reference's weight, score64 the weight with the correlation in FP64.
__global__ __launch_bounds__(256) void edge_weight_kernel(
    const uint32_t* __restrict__ efrom, const uint32_t* __restrict__ eto,
    const float* __restrict__ scomp, uint64_t ne, const float* __restrict__ unit,
    const float* __restrict__ raw, const double* __restrict__ colnorm, uint32_t S,
    double w_comp, double min_edge, double neg_thr, float* __restrict__ score,
    double* __restrict__ score64) {
  const uint64_t cnt = ne;
  for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < cnt;
       e += (uint64_t)gridDim.x * 256) {
    const uint32_t i = efrom[e], j = eto[e];
    double w, w64;
    if (S == 0) {
      w = w64 = (double)scomp[e] < min_edge ? 0.0 : (double)scomp[e];
    } else if (S == 1) {
      w = w64 = scomp[e];
    } else {
      const float* a = unit + (size_t)i * S;
      const float* b = unit + (size_t)j * S;
      const double corr = corr_ref_ptr(a, b, S);
      double corr64 = 0.0;
      for (uint32_t k = 0; k < S; ++k) corr64 += (double)a[k] * (double)b[k];
      const float* ra = raw + (size_t)i * S;
      const float* rb = raw + (size_t)j * S;
      double mn = 0.0, mx = 0.0;
      for (uint32_t k = 0; k < S; ++k) {
        double x = ra[k], y = rb[k];
        if (colnorm) {
          x *= colnorm[k];
          y *= colnorm[k];
        }
        if (x < y) {
          mn += x;
          mx += y;
        } else {
          mn += y;
          mx += x;
        }
      }
      const double wj = wjac_clamped(mn, mx);
      double pre;
      w = edge_weight_of(corr, wj, scomp[e], w_comp, min_edge, neg_thr, pre);
      w64 = edge_weight_of(corr64, wj, scomp[e], w_comp, min_edge, neg_thr, pre);
    }
    score[e] = (float)w;
    if (score64) score64[e] = w64;
  }
}

// The weights of the edges of a batch, each with its assembly's S
// (S <= 16*NC), w_comp and colnorm; unit and raw rows are zero-padded to
// 16*NC floats (128-byte aligned rows).  Edges whose FP64 weight lies within
// near_tol of a cut are appended to the near list (edge index, FP64 weight).
//
// A wave takes 64 consecutive edges, one per lane.  Edges are sorted by
// `from`, so the lanes' `from` rows mostly coincide and are read directly; the
// `to` rows are scattered, so the wave loads them together, 4*NC lanes per
// row (each load instruction reads whole rows), into an LDS area with an odd
// row stride (the lanes' reads of their own rows then fall in distinct banks).
constexpr uint32_t kEwWaves = 4;
template <int NC>
__device__ __forceinline__ void stage_rows(const float* __restrict__ base, uint32_t j, uint32_t ln,
                                           float* __restrict__ area) {
  constexpr int P = 4 * NC;          // float4 per row
  constexpr int RPI = 64 / P;        // rows per load instruction
  constexpr int STR = 16 * NC + 1;   // staged floats per row
  const uint32_t g = ln / P, q = ln % P;
#pragma unroll
  for (int r0 = 0; r0 < 64; r0 += RPI) {
    const uint32_t r = r0 + g;
    const uint32_t jr = (uint32_t)__shfl((int)j, (int)(r < 64 ? r : 0));
    if (g < (uint32_t)RPI && r < 64) {
      const float4 v = reinterpret_cast<const float4*>(base + (size_t)jr * 16 * NC)[q];
      float* dst = area + r * STR + 4 * q;
      dst[0] = v.x;
      dst[1] = v.y;
      dst[2] = v.z;
      dst[3] = v.w;
    }
  }
  __builtin_amdgcn_wave_barrier();  // rows written by other lanes (LDS keeps a wave's order)
}

template <int NC>
__global__ __launch_bounds__(64 * kEwWaves) void edge_weight_pad_kernel(
    const uint32_t* __restrict__ efrom, const uint32_t* __restrict__ eto,
    const float* __restrict__ scomp, uint64_t cap, const uint64_t* __restrict__ ne_dev,
    const float* __restrict__ unitp, const float* __restrict__ rawp,
    const double* __restrict__ colnorm_all, const AsmDesc* __restrict__ desc,
    const uint16_t* __restrict__ asm_of, double min_edge, double neg_thr, double near_tol,
    float* __restrict__ score, unsigned long long* __restrict__ near_cnt, uint64_t near_cap,
    uint64_t* __restrict__ near_edge, double* __restrict__ near_w64) {
  constexpr int V = 4 * NC;          // float4 per row
  constexpr int STR = 16 * NC + 1;
  __shared__ float stage[kEwWaves][64 * STR];
  const uint32_t ln = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  float* __restrict__ area = stage[wv];
  const float* __restrict__ mine = area + ln * STR;
  const uint64_t cnt = min(*ne_dev, cap);
  const uint64_t wstride = (uint64_t)gridDim.x * 64 * kEwWaves;
  for (uint64_t e0 = ((uint64_t)blockIdx.x * kEwWaves + wv) * 64; e0 < cnt; e0 += wstride) {  // wave-uniform
    const uint64_t e = e0 + ln;
    const bool live = e < cnt;
    const uint32_t i = live ? efrom[e] : 0, j = live ? eto[e] : 0;
    const AsmDesc& ad = desc[asm_of[i]];
    const uint32_t S = ad.S;
    const double w_comp = ad.w_comp;
    // unit rows: own `from` row in registers, the `to` rows staged
    stage_rows<NC>(unitp, j, ln, area);
    float a[4 * V];
    {
      const float4* ua = reinterpret_cast<const float4*>(unitp + (size_t)i * 16 * NC);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const float4 x = ua[v];
        a[4 * v] = x.x, a[4 * v + 1] = x.y, a[4 * v + 2] = x.z, a[4 * v + 3] = x.w;
      }
    }
    double corr64;
    const double corr = corr_ref_regs<4 * V>(a, mine, S, corr64);
    const bool cut = !finite64(corr) || (neg_thr > -0.99 && corr < neg_thr);
    const bool cut64 = !finite64(corr64) || (neg_thr > -0.99 && corr64 < neg_thr);
    const bool depth = live && S > 1 && !(cut && cut64);  // a weighted Jaccard is needed
    __builtin_amdgcn_wave_barrier();  // every lane has read its staged unit row
    const bool any_depth = __ballot(depth) != 0;
    if (any_depth) stage_rows<NC>(rawp, j, ln, area);  // wave-uniform
    double w = 0.0;
    bool near = false;
    double w64 = 0.0;
    if (depth) {
      const double* __restrict__ colnorm = colnorm_all + ad.colnorm0;
      const float4* ra = reinterpret_cast<const float4*>(rawp + (size_t)i * 16 * NC);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const float4 x = ra[v];
        a[4 * v] = x.x, a[4 * v + 1] = x.y, a[4 * v + 2] = x.z, a[4 * v + 3] = x.w;
      }
      double mn = 0.0, mx = 0.0;
#pragma unroll
      for (int k = 0; k < 4 * V; ++k) {
        if ((uint32_t)k < S) {
          double x = a[k], y = mine[k];
          if (colnorm) {
            x *= colnorm[k];
            y *= colnorm[k];
          }
          if (x < y) {
            mn += x;
            mx += y;
          } else {
            mn += y;
            mx += x;
          }
        }
      }
      const double wj = wjac_clamped(mn, mx);
      const double sc = (double)scomp[e];
      double pre, pre64;
      w = edge_weight_of(corr, wj, sc, w_comp, min_edge, neg_thr, pre);
      w64 = edge_weight_of(corr64, wj, sc, w_comp, min_edge, neg_thr, pre64);
      near = near_cut(corr64, pre64, min_edge, neg_thr, near_tol);
    } else if (live && S > 1) {  // both correlations cut the edge
      near = neg_thr > -0.99 && fabs(corr64 - neg_thr) <= near_tol;
    } else if (live) {  // S <= 1: the composition similarity alone (exact)
      const double sc = (double)scomp[e];
      w = S == 0 && sc < min_edge ? 0.0 : sc;
    }
    if (live) score[e] = (float)w;
    if (near) {
      const unsigned long long t = atomicAdd(near_cnt, 1ull);
      if (t < near_cap) {
        near_edge[t] = e;
        near_w64[t] = w64;
      }
    }
    __builtin_amdgcn_wave_barrier();  // the staged rows are read before the next edges' load
  }
}

// ---------------------------------------------------------------------------
// Build scalars on the device, so a graph build needs no host round trip
// between its stages.
//
// baseline_tables_kernel: b0 = the median of (float)(cnt * (1/m)) over the
// sampled pairs (src/rabbitbin.cpp:1340-1369); cnt -> float is monotone, so
// the median value is that of the median count, found in a count histogram.
// Then the sv tables of pmh_sv_from_count (src/rabbitbin.cpp:1297-1305) in
// FP64 (IEEE division and no contraction, as the host build) and the level
// of each count (dense rank of the distinct float similarities).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void baseline_tables_kernel(const uint32_t* __restrict__ pcnt_all,
                                                              const AsmDesc* __restrict__ desc,
                                                              uint32_t m,
                                                              BuildScalars* __restrict__ bs_all,
                                                              double* __restrict__ svd_all,
                                                              float* __restrict__ svf_all,
                                                              uint16_t* __restrict__ level_all) {
  __shared__ uint32_t hist[513];
  __shared__ float sf[513];
  __shared__ double b0s;
  const uint32_t tid = threadIdx.x;
  const uint32_t a = blockIdx.x;  // one assembly per block
  const uint32_t np = desc[a].np;
  const uint32_t* __restrict__ pcnt = pcnt_all + desc[a].pair0;
  BuildScalars* __restrict__ bs = bs_all + a;
  double* __restrict__ svd = svd_all + (size_t)a * (m + 1);
  float* __restrict__ svf = svf_all + (size_t)a * (m + 1);
  uint16_t* __restrict__ level = level_all + (size_t)a * (m + 1);
  for (uint32_t c = tid; c <= m; c += 256) hist[c] = 0;
  __syncthreads();
  for (uint32_t t = tid; t < np; t += 256) atomicAdd(&hist[pcnt[t]], 1u);
  __syncthreads();
  const double inv_m = 1.0 / (double)m;
  if (tid == 0) {
    double b0 = 0.0;
    if (np > 0) {
      const uint32_t k = np / 2;  // vals[size / 2] after nth_element
      uint32_t cum = 0, c = 0;
      for (; c < m; ++c) {
        cum += hist[c];
        if (cum > k) break;
      }
      b0 = (double)(float)((double)c * inv_m);
    }
    const double given = desc[a].b0_in;
    if (given == given) b0 = given;  // the caller's b0
    b0s = b0;
    bs->b0 = b0;
  }
  __syncthreads();
  const double b0 = b0s;
  const bool rescale = b0 > 0.0 && b0 < 1.0;
  const double inv1mb0 = rescale ? 1.0 / (1.0 - b0) : 1.0;
  for (uint32_t c = tid; c <= m; c += 256) {
    double j = (double)c * inv_m;
    if (rescale) {
      j = (j - b0) * inv1mb0;
      if (j < 0.0) j = 0.0;
    }
    if (j > 1.0 - 1e-6) j = 1.0 - 1e-6;
    svd[c] = j;
    svf[c] = (float)j;
    sf[c] = (float)j;
  }
  __syncthreads();
  if (tid == 0) {
    uint16_t lv = 0;
    for (uint32_t c = 0; c <= m; ++c) {
      if (c > 0 && sf[c] != sf[c - 1]) ++lv;
      level[c] = lv;
    }
    bs->nlevels = (uint32_t)lv + 1;
  }
}

// calibrate_kernel: simCutoff from the per-row maximum counts, as
// calibrate_sim_cutoff(coverage, true) (regime 1, src/impl/rb_graph.cpp:
// 1362-1384) or the 10 x 2500 sampled calibration of gen_fused_calib_graph
// (regime 2, :502-525, 1243-1254).  The contig draws of both depend only on n
// and the srand() seed and are done on the host: `rnd` holds the rand()
// values that follow them (consumed by calib_converge, :390-415) and, in
// regime 2, members[round_off[r] .. round_off[r+1]) are the contigs whose
// lowest sampled round is r (their maxima count in round r; the other slots
// of a round hold 0).  maxsim values are svd[cnt] (regime 1) or (double)
// svf[cnt] (regime 2) and increase with cnt, so the number of contigs at or
// above a cutoff is a suffix sum of a count histogram.  Also writes cmin, the
// first count whose float similarity exceeds simCutoff / 1000.
// Block-wide exclusive scan (256 threads; defined with the scan kernels).
__device__ uint64_t block_scan_excl_u64(uint64_t v, uint64_t* wave_tot, uint64_t* total);

// calib_converge over the suffix histogram of one round: cov(p) = (contigs
// whose maxsim >= p / 1000) / denom = suf[lo_at[p]] / denom, with lo_at[p] the
// first count whose maxsim reaches p / 1000 (m + 1 when none does:
// suf[m + 1] = 0).  fl(c / denom) increases with c, so cov >= coverage exactly
// when c >= cstar (the first such c): the walk compares integers and divides
// only where it stops.  rmod[i] = rand() value i % 3.
__device__ uint32_t calib_converge_dev(const uint32_t* suf, const uint16_t* lo_at, double denom,
                                       uint32_t cstar, double coverage, const uint8_t* rmod, uint32_t nrnd,
                                       uint32_t& ri) {
  uint32_t p = 999, pp = 1000, pc = 0;
  while (p > 50) {
    const uint32_t c = suf[lo_at[p]];
    if (c >= cstar) {
      const double cov = (double)c / denom, pcov = (double)pc / denom;
      if (cov - coverage > coverage - pcov) p = pp;
      break;
    }
    pp = p;
    pc = c;
    const int r = ri < nrnd ? (int)rmod[ri] : 0;
    ++ri;
    p -= (uint32_t)(p > 990 ? r + 1 : p > 900 ? r + 3 : r + 9);
  }
  return p;
}

constexpr uint32_t kCalibRounds = 10;
constexpr uint32_t kCalibBins = 520;  // counts 0..m+1 (m <= 512), per round
static_assert(kCalibRounds * kCalibBins == kCalibHistWords, "calibration histogram scratch");

// Every round's count histogram of every assembly (hist: kCalibRounds x
// kCalibBins per assembly, zeroed before), one thread per sampled contig:
// regime 1 every row of the assembly (n > 1), regime 2 members[t] for t in
// [round_off[0], round_off[10]), round r holding [round_off[r],
// round_off[r + 1]) (members are global rows).  Spread over the device, the
// contigs' maxima are gathered in parallel instead of by one workgroup.
__global__ __launch_bounds__(256) void calib_hist_kernel(const uint32_t* __restrict__ rowmax_all,
                                                         const AsmDesc* __restrict__ desc,
                                                         const uint32_t* __restrict__ members,
                                                         const uint32_t* __restrict__ round_off_all,
                                                         uint32_t* __restrict__ hist_all) {
  const uint32_t a = blockIdx.y;
  const AsmDesc d = desc[a];
  const uint32_t* __restrict__ round_off = round_off_all + (size_t)a * 11;
  const bool small = d.regime == 1;
  const uint32_t t_beg = small ? 0u : round_off[0];
  const uint32_t t_end = small ? (d.n > 1 ? d.n : 0u) : round_off[kCalibRounds];
  const uint32_t t = t_beg + blockIdx.x * 256 + threadIdx.x;
  if (t >= t_end) return;
  const uint32_t c = small ? rowmax_all[d.row0 + t] : rowmax_all[members[t]];
  uint32_t r = 0;
  if (!small)
    while (t >= round_off[r + 1]) ++r;
  atomicAdd(&hist_all[((size_t)a * kCalibRounds + r) * kCalibBins + c], 1u);
}

// One workgroup per assembly.  Each wave turns its rounds' histograms
// (calib_hist_kernel) into suffix sums, and thread 0 then runs the rounds'
// convergence walks back to back (each reads the rand() values where the
// previous one stopped); no workgroup barrier between rounds.
__global__ __launch_bounds__(256) void calibrate_kernel(
    const uint32_t* __restrict__ rowmax_all, const AsmDesc* __restrict__ desc, uint32_t m,
    const double* __restrict__ svd_all, const float* __restrict__ svf_all,
    const int32_t* __restrict__ rnd_all, const uint32_t* __restrict__ members,
    const uint32_t* __restrict__ round_off_all, const uint32_t* __restrict__ hist_all, double coverage,
    BuildScalars* __restrict__ bs_all) {
  const uint32_t a = blockIdx.x;  // one assembly per block
  const uint32_t n = desc[a].n;
  const int regime = (int)desc[a].regime;
  const uint32_t* __restrict__ rowmax = rowmax_all + desc[a].row0;  // regime 1 rows
  const double* __restrict__ svd = svd_all + (size_t)a * (m + 1);
  const float* __restrict__ svf = svf_all + (size_t)a * (m + 1);
  const uint32_t nrnd = kCalibRandDev;
  const uint32_t* __restrict__ round_off = round_off_all + (size_t)a * 11;
  BuildScalars* __restrict__ bs = bs_all + a;
  __shared__ uint32_t hist[kCalibRounds][kCalibBins];
  __shared__ uint16_t lo_at[1000];
  __shared__ double val[513];
  // the rand() values % 3, read one after another by the serial convergence walks
  __shared__ uint8_t rmod[kCalibRandDev];
  __shared__ uint32_t ri_s, cmin_s;
  __shared__ double sum_p;
  const uint32_t tid = threadIdx.x, ln = tid & 63u, wv = tid >> 6;
  const uint32_t rounds = regime == 1 ? 1 : kCalibRounds;
  for (uint32_t c = tid; c <= m; c += 256) val[c] = regime == 1 ? svd[c] : (double)svf[c];
  for (uint32_t q = tid; q < kCalibRandDev; q += 256) rmod[q] = (uint8_t)(rnd_all[(size_t)a * kCalibRandDev + q] % 3);
  for (uint32_t c = tid; c < rounds * kCalibBins; c += 256)
    (&hist[0][0])[c] = hist_all[(size_t)a * kCalibRounds * kCalibBins + c];
  // lo_at[p]: the first count whose maxsim reaches p / 1000 (val increases
  // with the count)
  for (uint32_t p = tid; p < 1000; p += 256) {
    const double cut = (double)p / 1000.;
    uint32_t lo = 0, hi = m + 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (val[mid] >= cut)
        hi = mid;
      else
        lo = mid + 1;
    }
    lo_at[p] = (uint16_t)lo;
  }
  __syncthreads();
  // suffix sums, one wave per round: hist[r][c] <- contigs with count >= c
  // (lane l takes bins 9l .. 9l+8; counts run to m <= 512, bin m + 1 stays 0)
  for (uint32_t r = wv; r < rounds; r += 4) {
    uint32_t* __restrict__ h = hist[r];
    uint32_t loc[9], part = 0;
#pragma unroll
    for (uint32_t k = 0; k < 9; ++k) {
      const uint32_t c = 9 * ln + k;
      loc[k] = c <= m + 1 ? h[c] : 0u;
      part += loc[k];
    }
    uint32_t incl = part;  // lanes >= ln
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_down((int)incl, d);
      if (ln + d < 64) incl += t;
    }
    uint32_t run = incl - part;  // bins above this lane's
#pragma unroll
    for (int k = 8; k >= 0; --k) {
      const uint32_t c = 9 * ln + (uint32_t)k;
      run += loc[k];
      if (c <= m + 1) h[c] = run;
    }
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t ri = 0;
    double sp = 0.0;
    const double denom = regime == 1 ? (double)n : 2500.0;
    // cstar: the first count c with fl(c / denom) >= coverage (bisection over
    // the counts a round can hold; maxc + 1 when none reaches it)
    const uint32_t maxc = regime == 1 ? n : 2500u;
    uint32_t lo = 0, hi = maxc + 1;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if ((double)mid / denom >= coverage)
        hi = mid;
      else
        lo = mid + 1;
    }
    const uint32_t cstar = lo;
    for (uint32_t r = 0; r < rounds; ++r) {
      const uint32_t p = calib_converge_dev(hist[r], lo_at, denom, cstar, coverage, rmod, nrnd, ri);
      if (regime == 1) {
        sp = (double)p;
      } else {
        double mp = (double)p;
        if (mp < 51.0) mp = 50.0;
        sp += mp;
        if (r == 1 && sp / 2 < 51.0) {
          sp = 50.0 * 10.0;
          break;
        }
      }
    }
    ri_s = ri;
    sum_p = sp;
  }
  if (tid == 0) cmin_s = m + 1;
  __syncthreads();
  const double sim_cutoff = regime == 1 ? sum_p : sum_p / 10.0;
  const double cutoff = sim_cutoff / 1000.;
  // cmin: the first count whose float similarity exceeds the cutoff
  for (uint32_t c = tid; c <= m; c += 256)
    if ((double)svf[c] > cutoff) atomicMin(&cmin_s, c);
  __syncthreads();
  if (tid == 0) {
    bs->sim_cutoff = sim_cutoff;
    bs->cmin = cmin_s;
    bs->rand_used = ri_s;
  }
}

// Exclusive prefix sum u32 -> u64 in three steps: per-block sums (1024
// elements per block), a one-block scan of the block sums, then a per-block
// scan with the block's offset.  out has n + 1 entries (out[n] = total).
__device__ uint64_t block_scan_excl_u64(uint64_t v, uint64_t* wave_tot, uint64_t* total) {
  const uint32_t tid = threadIdx.x, ln = tid & 63u;
  uint64_t incl = v;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint64_t t = __shfl_up(incl, d);
    if (ln >= d) incl += t;
  }
  if (ln == 63) wave_tot[tid >> 6] = incl;
  __syncthreads();
  uint64_t excl = incl - v;
  for (uint32_t w = 0; w < (tid >> 6); ++w) excl += wave_tot[w];
  if (tid == 255) *total = excl + v;
  __syncthreads();
  return excl;
}

__global__ __launch_bounds__(256) void scan_block_sums_kernel(const uint32_t* __restrict__ in,
                                                              uint32_t n,
                                                              uint64_t* __restrict__ sums) {
  __shared__ uint64_t wt[4], tot;
  const uint64_t base = (uint64_t)blockIdx.x * 1024;
  uint64_t v = 0;
  for (uint32_t q = 0; q < 4; ++q) {
    const uint64_t idx = base + threadIdx.x * 4 + q;
    if (idx < n) v += in[idx];
  }
  block_scan_excl_u64(v, wt, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void scan_sums_kernel(uint64_t* __restrict__ sums, uint32_t nb) {
  __shared__ uint64_t wt[4], tot;
  uint64_t carry = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
    const uint32_t b = b0 + threadIdx.x;
    const uint64_t v = b < nb ? sums[b] : 0;
    const uint64_t ex = block_scan_excl_u64(v, wt, &tot);
    if (b < nb) sums[b] = carry + ex;
    carry += tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void scan_final_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                                         const uint64_t* __restrict__ sums,
                                                         uint64_t* __restrict__ out) {
  __shared__ uint64_t wt[4], tot;
  const uint64_t base = (uint64_t)blockIdx.x * 1024;
  uint32_t x[4];
  uint64_t v = 0;
  for (uint32_t q = 0; q < 4; ++q) {
    const uint64_t idx = base + threadIdx.x * 4 + q;
    x[q] = idx < n ? in[idx] : 0u;
    v += x[q];
  }
  uint64_t run = sums[blockIdx.x] + block_scan_excl_u64(v, wt, &tot);
  for (uint32_t q = 0; q < 4; ++q) {
    const uint64_t idx = base + threadIdx.x * 4 + q;
    if (idx <= n) out[idx] = run;  // idx == n writes the grand total
    run += x[q];
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------
hipError_t launch_sketch(const uint8_t* seqs, const uint64_t* offs, uint32_t n, uint32_t m,
                         uint32_t W, const double* sT, const uint8_t* sX, const uint8_t* cls,
                         const uint8_t* codes, uint8_t* wx, uint32_t mpad, uint32_t* planes,
                         const uint32_t* out_row, hipStream_t st) {
  constexpr int CPB = 4;
  const uint32_t blocks = (n + CPB - 1) / CPB;
  if (blocks == 0) return hipSuccess;
  if (W > 16) return hipErrorInvalidValue;  // the kernel holds the winners of at most 512 registers
  hipLaunchKernelGGL(sketch_kernel<CPB>, dim3(blocks), dim3(256), 0, st, seqs, offs, n, m, W, sT,
                     sX, cls, codes, wx, mpad, planes, out_row);
  return hipGetLastError();
}

hipError_t launch_planes_from_classes(const uint8_t* cls, uint32_t n, uint32_t m, uint32_t mpad,
                                      uint32_t W, uint32_t row0, uint32_t* planes, hipStream_t st) {
  const uint64_t waves = (uint64_t)n * ((32u * W + 63u) / 64u);
  if (waves == 0) return hipSuccess;
  hipLaunchKernelGGL(planes_from_classes_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, cls, n,
                     m, mpad, W, row0, planes);
  return hipGetLastError();
}

hipError_t launch_list_counts(const uint32_t* planes, uint32_t W, const uint32_t* pi,
                              const uint32_t* pj, uint32_t np, uint32_t* out, hipStream_t st) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(list_counts_kernel, dim3((np + 255) / 256), dim3(256), 0, st, planes, W, pi,
                     pj, np, out);
  return hipGetLastError();
}

hipError_t launch_win16_counts(const uint16_t* win16, uint32_t m, const uint32_t* pi,
                               const uint32_t* pj, uint32_t np, uint32_t* out, hipStream_t st) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(win16_counts_kernel, dim3((np + 255) / 256), dim3(256), 0, st, win16, m, pi, pj,
                     np, out);
  return hipGetLastError();
}

void dense_work_items(const std::vector<AsmDesc>& desc, uint32_t rows_per_item,
                      std::vector<uint32_t>& items) {
  items.clear();
  for (uint32_t a = 0; a < (uint32_t)desc.size(); ++a) {
    if (!desc[a].dense) continue;
    const uint32_t n = desc[a].n;
    for (uint32_t rb = 0; rb * rows_per_item < n; ++rb)
      for (uint32_t cb = 0; cb * 256 < n; ++cb) {
        const uint32_t jl = std::min(n, cb * 256 + 256) - 1;  // the block's last column
        if (jl <= rb * rows_per_item || jl < desc[a].jmin) continue;  // no pair r < j, j >= jmin
        items.push_back(a);
        items.push_back(cb);
        items.push_back(rb);
        items.push_back(rows_per_item);
      }
  }
}

hipError_t launch_dense_counts(const uint32_t* planes, uint32_t W, const AsmDesc* desc,
                               const uint32_t* items, uint32_t nitems, uint16_t* mat,
                               hipStream_t st) {
  if (nitems == 0) return hipSuccess;
  const uint4* it = reinterpret_cast<const uint4*>(items);
  const dim3 grid(nitems);
  switch (W) {
    case 16: hipLaunchKernelGGL(dense_counts_kernel<16>, grid, dim3(256), 0, st, planes, desc, it, mat); break;
    case 8: hipLaunchKernelGGL(dense_counts_kernel<8>, grid, dim3(256), 0, st, planes, desc, it, mat); break;
    case 4: hipLaunchKernelGGL(dense_counts_kernel<4>, grid, dim3(256), 0, st, planes, desc, it, mat); break;
    case 2: hipLaunchKernelGGL(dense_counts_kernel<2>, grid, dim3(256), 0, st, planes, desc, it, mat); break;
    case 1: hipLaunchKernelGGL(dense_counts_kernel<1>, grid, dim3(256), 0, st, planes, desc, it, mat); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

void gate_work_items(uint32_t a, uint32_t n, uint32_t jmin, uint32_t tiles_per_item,
                     std::vector<uint32_t>& full, std::vector<uint32_t>& part) {
  // (a, bi << 16 | bj0, tile count, jmin); items with the full tile count go first
  const uint32_t T = (n + 255) / 256;
  for (uint32_t bi = 0; bi < T; ++bi)
    for (uint32_t bj = std::max(bi, jmin / 256); bj < T; bj += tiles_per_item) {
      const uint32_t cnt = std::min(tiles_per_item, T - bj);
      auto& dst = cnt == tiles_per_item ? full : part;
      dst.push_back(a);
      dst.push_back((bi << 16) | bj);
      dst.push_back(cnt);
      dst.push_back(jmin);
    }
}

hipError_t launch_gate_mfma(const int8_t* q, uint32_t kpad, const AsmDesc* desc,
                            const uint32_t* perm, const uint32_t* items, uint32_t nitems,
                            uint64_t* cand, unsigned long long* ncand, uint64_t cap,
                            hipStream_t st) {
  if (nitems == 0) return hipSuccess;
  const uint4* it = reinterpret_cast<const uint4*>(items);
  switch (kpad / 32) {
    case 1: hipLaunchKernelGGL(gate_mfma_kernel<1>, dim3(nitems), dim3(256), 0, st, q, desc, perm, it, cand, ncand, cap); break;
    case 2: hipLaunchKernelGGL(gate_mfma_kernel<2>, dim3(nitems), dim3(256), 0, st, q, desc, perm, it, cand, ncand, cap); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_gate_compare(uint64_t* cand, uint64_t cap, const unsigned long long* ncand,
                               const float* unit, uint32_t NC, const uint8_t* anynz,
                               const AsmDesc* desc, const uint16_t* asm_of, uint32_t n_asm,
                               const uint32_t* planes, uint32_t W, uint32_t* deg, uint32_t* slot,
                               hipStream_t st) {
  if (cap == 0) return hipSuccess;
  if (cap > 0x7FFFFFFFull) return hipErrorInvalidValue;  // slot indices 2t are 32-bit
  // a multiple of the XCD count (xcd_segment)
  const dim3 grid((unsigned)std::max<uint64_t>(kXcds, std::min<uint64_t>((cap * 8 + 255) / 256, 65536) & ~7ull));
#define RBG_GC(NCV, WV)                                                                            \
  hipLaunchKernelGGL((gate_compare_kernel<NCV, WV>), grid, dim3(256), 0, st, cand, cap, ncand, unit, \
                     anynz, desc, asm_of, n_asm, planes, deg, slot)
#define RBG_GC_W(NCV)                   \
  switch (W) {                          \
    case 16: RBG_GC(NCV, 16); break;    \
    case 8: RBG_GC(NCV, 8); break;      \
    case 4: RBG_GC(NCV, 4); break;      \
    case 2: RBG_GC(NCV, 2); break;      \
    case 1: RBG_GC(NCV, 1); break;      \
    default: return hipErrorInvalidValue; \
  }
  switch (NC) {
    case 1: RBG_GC_W(1); break;
    case 2: RBG_GC_W(2); break;
    case 3: RBG_GC_W(3); break;
    case 4: RBG_GC_W(4); break;
    default: return hipErrorInvalidValue;
  }
#undef RBG_GC_W
#undef RBG_GC
  return hipGetLastError();
}

hipError_t launch_scatter(const uint64_t* cand, uint64_t cap, const unsigned long long* ncand,
                          const uint32_t* slot, const uint64_t* off, uint64_t* list, hipStream_t st) {
  if (cap == 0) return hipSuccess;
  const dim3 grid((unsigned)std::min<uint64_t>((cap + 255) / 256, 65536));
  hipLaunchKernelGGL(scatter_kernel, grid, dim3(256), 0, st, cand, cap, ncand,
                     reinterpret_cast<const uint2*>(slot), off, list);
  return hipGetLastError();
}

hipError_t launch_merge_deg(const MergeDesc* md, uint32_t nmd, uint32_t max_old, uint32_t* deg,
                            hipStream_t st) {
  if (nmd == 0 || max_old == 0) return hipSuccess;
  hipLaunchKernelGGL(merge_deg_kernel, dim3((max_old + 255) / 256, nmd), dim3(256), 0, st, md, deg);
  return hipGetLastError();
}

hipError_t launch_merge_lists(const MergeDesc* md, uint32_t nmd, uint32_t max_old, const uint64_t* off,
                              uint64_t* list, hipStream_t st) {
  if (nmd == 0 || max_old == 0) return hipSuccess;
  hipLaunchKernelGGL(merge_lists_kernel, dim3((max_old + 3) / 4, nmd), dim3(256), 0, st, md, off,
                     list);
  return hipGetLastError();
}

hipError_t launch_extract_lists(const uint64_t* off, const uint64_t* list, uint32_t row0, uint32_t n,
                                uint64_t* loff, uint64_t* dst, hipStream_t st) {
  hipLaunchKernelGGL(extract_lists_kernel, dim3((n + 1 + 3) / 4), dim3(256), 0, st, off, list, row0,
                     n, loff, dst);
  return hipGetLastError();
}

hipError_t launch_select(const uint64_t* list, const uint64_t* off, const uint16_t* mat,
                         const uint8_t* anynz, const AsmDesc* desc, const uint16_t* asm_of,
                         uint32_t nrows, const uint16_t* level, uint32_t m, uint32_t K,
                         uint32_t* topk_id, uint16_t* topk_cnt, uint32_t* topk_len,
                         uint32_t* rowmax, uint32_t* fence, uint32_t FW, uint32_t* queue,
                         uint32_t nstatic, hipStream_t st) {
  if (nrows == 0) return hipSuccess;
  if (m + 1 > 1024 || K > 1024 || nstatic > nrows) return hipErrorInvalidValue;
  uint32_t* qcount = queue + nrows;  // queued long lists, rows taken by the workgroup pass
  const hipError_t e = launch_fill_u32(qcount, 2, 0u, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(select_wave_kernel, dim3((nrows + 3) / 4), dim3(256), 0, st, list, off, desc, asm_of,
                     level, m, nrows, K, topk_id, topk_cnt, topk_len, rowmax, fence, FW, queue, nstatic, qcount);
  // the listed dense rows, then the queued long lists (their number is on
  // the device)
  if (nstatic)
    hipLaunchKernelGGL(select_topk_kernel, dim3(nstatic), dim3(256), 0, st, queue, list, off, mat, anynz, desc,
                       asm_of, level, m, K, topk_id, topk_cnt, topk_len, rowmax, fence, FW);
  hipLaunchKernelGGL(select_long_kernel, dim3(std::min<uint32_t>(nrows, 2048)), dim3(256), 0, st, queue, nstatic,
                     qcount, list, off, mat, anynz, desc, asm_of, level, m, K, topk_id, topk_cnt, topk_len,
                     rowmax, fence, FW);
  return hipGetLastError();
}

hipError_t launch_emit_mark(const uint32_t* topk_id, const uint16_t* topk_cnt,
                            const uint32_t* topk_len, const uint32_t* fence, uint32_t FW,
                            uint32_t n, uint32_t K, const BuildScalars* bs,
                            const AsmDesc* desc, const uint16_t* asm_of, const uint32_t* order,
                            uint64_t* keep, uint32_t* row_edges, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(emit_mark_kernel, dim3((n + 3) / 4), dim3(256), 0, st, topk_id, topk_cnt,
                     topk_len, fence, FW, n, K, bs, desc, asm_of, order, (K + 63) / 64, keep,
                     row_edges);
  return hipGetLastError();
}

hipError_t launch_emit_write(const uint32_t* topk_id, const uint16_t* topk_cnt,
                             const uint32_t* topk_len, uint32_t n, uint32_t K, const uint64_t* keep,
                             const uint64_t* row_off, const uint16_t* asm_of, uint32_t m,
                             const float* svf, uint32_t* efrom, uint32_t* eto, uint16_t* ecnt,
                             float* escomp, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(emit_write_kernel, dim3((n + 3) / 4), dim3(256), 0, st, topk_id, topk_cnt,
                     topk_len, n, K, (K + 63) / 64, keep, row_off, asm_of, m, svf, efrom, eto, ecnt,
                     escomp);
  return hipGetLastError();
}

hipError_t launch_edge_weights_batch(const uint32_t* efrom, const uint32_t* eto,
                                     const float* scomp, uint64_t cap, const uint64_t* ne,
                                     const float* unit, const float* raw, uint32_t NC,
                                     const double* colnorm, const AsmDesc* desc,
                                     const uint16_t* asm_of, double min_edge, double neg_thr,
                                     double near_tol, float* score, unsigned long long* near_cnt,
                                     uint64_t near_cap, uint64_t* near_edge, double* near_w64,
                                     hipStream_t st) {
  if (cap == 0) return hipSuccess;
  const dim3 grid((unsigned)std::min<uint64_t>((cap + 64 * kEwWaves - 1) / (64 * kEwWaves), 65536));
#define RBG_EW(NCV)                                                                                   \
  hipLaunchKernelGGL(edge_weight_pad_kernel<NCV>, grid, dim3(64 * kEwWaves), 0, st, efrom, eto, scomp, \
                     cap, ne, unit, raw, colnorm, desc, asm_of, min_edge, neg_thr, near_tol, score,   \
                     near_cnt, near_cap, near_edge, near_w64)
  switch (NC) {
    case 1: RBG_EW(1); break;
    case 2: RBG_EW(2); break;
    case 3: RBG_EW(3); break;
    case 4: RBG_EW(4); break;
    default: return hipErrorInvalidValue;
  }
#undef RBG_EW
  return hipGetLastError();
}

hipError_t launch_edge_weights(const uint32_t* efrom, const uint32_t* eto, const float* scomp,
                               uint64_t ne, const float* unit, const float* raw,
                               const double* colnorm, uint32_t S, double w_comp, double min_edge,
                               double neg_thr, float* score, double* score64, hipStream_t st) {
  if (ne == 0) return hipSuccess;
  hipLaunchKernelGGL(edge_weight_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, efrom,
                     eto, scomp, ne, unit, raw, colnorm, S, w_comp, min_edge, neg_thr, score, score64);
  return hipGetLastError();
}

hipError_t exclusive_scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint32_t n, void* tmp,
                                     size_t& tmp_bytes, hipStream_t st) {
  // n + 1 outputs; blocks cover 1024 inputs each (the last block also writes out[n]).
  const uint32_t nb = n / 1024 + 1;
  if (tmp == nullptr) {
    tmp_bytes = (size_t)nb * sizeof(uint64_t);
    return hipSuccess;
  }
  uint64_t* sums = static_cast<uint64_t*>(tmp);
  hipLaunchKernelGGL(scan_block_sums_kernel, dim3(nb), dim3(256), 0, st, in, n, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(256), 0, st, sums, nb);
  hipLaunchKernelGGL(scan_final_kernel, dim3(nb), dim3(256), 0, st, in, n, sums, out);
  return hipGetLastError();
}

hipError_t launch_baseline_tables(const uint32_t* pcnt, const AsmDesc* desc, uint32_t nasm,
                                  uint32_t m, BuildScalars* bs, double* svd, float* svf,
                                  uint16_t* level, hipStream_t st) {
  if (m < 2 || m > 512) return hipErrorInvalidValue;
  if (nasm == 0) return hipSuccess;
  hipLaunchKernelGGL(baseline_tables_kernel, dim3(nasm), dim3(256), 0, st, pcnt, desc, m, bs, svd,
                     svf, level);
  return hipGetLastError();
}

hipError_t launch_calibrate(const uint32_t* rowmax, const AsmDesc* desc, uint32_t nasm, uint32_t m,
                            const double* svd, const float* svf, const int32_t* rnd,
                            const uint32_t* members, const uint32_t* round_off, uint32_t max_items,
                            uint32_t* hist, double coverage, BuildScalars* bs, hipStream_t st) {
  if (m < 2 || m > 512) return hipErrorInvalidValue;
  if (nasm == 0) return hipSuccess;
  hipError_t e = launch_fill_u32(hist, (uint64_t)nasm * kCalibRounds * kCalibBins, 0u, st);
  if (e != hipSuccess) return e;
  if (max_items)
    hipLaunchKernelGGL(calib_hist_kernel, dim3((max_items + 255) / 256, nasm), dim3(256), 0, st, rowmax, desc,
                       members, round_off, hist);
  hipLaunchKernelGGL(calibrate_kernel, dim3(nasm), dim3(256), 0, st, rowmax, desc, m, svd, svf, rnd,
                     members, round_off, hist, coverage, bs);
  return hipGetLastError();
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t count, uint32_t value) {
  for (uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x; t < count;
       t += (uint64_t)gridDim.x * 256)
    p[t] = value;
}

__global__ void gather_bounds_kernel(const uint64_t* __restrict__ row_off,
                                     const uint64_t* __restrict__ off, const AsmDesc* __restrict__ desc,
                                     uint32_t nasm, uint64_t* __restrict__ out) {
  const uint32_t a = blockIdx.x * 64 + threadIdx.x;
  if (a >= nasm) return;
  const uint32_t r0 = desc[a].row0, r1 = r0 + desc[a].n;
  out[4 * a] = row_off[r0];
  out[4 * a + 1] = row_off[r1];
  out[4 * a + 2] = off ? off[r0] : 0;
  out[4 * a + 3] = off ? off[r1] : 0;
}

hipError_t launch_gather_bounds(const uint64_t* row_off, const uint64_t* off, const AsmDesc* desc,
                                uint32_t nasm, uint64_t* out, hipStream_t st) {
  if (nasm == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_bounds_kernel, dim3((nasm + 63) / 64), dim3(64), 0, st, row_off, off,
                     desc, nasm, out);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t count, uint32_t value, hipStream_t st) {
  if (count == 0) return hipSuccess;
  const dim3 grid((unsigned)std::min<uint64_t>((count + 255) / 256, 65536));
  hipLaunchKernelGGL(fill_u32_kernel, grid, dim3(256), 0, st, p, count, value);
  return hipGetLastError();
}

}  // namespace rbg
