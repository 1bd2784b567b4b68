// Ingest kernels' shared types and launchers (internal to librbg).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbg {

// Record classes of the FASTA parse.
enum : uint32_t { kFaDropped = 0, kFaTiny = 1, kFaSmall = 2, kFaLarge = 3 };

// One FASTA record: name [name_beg, name_beg + name_len) and body
// [body_beg, body_end) in the file; seq_len kept bases; single: the body's
// sequence is its first line.
struct FaRecord {
  uint64_t name_beg;
  uint64_t body_beg;
  uint64_t body_end;
  uint64_t seq_len;
  uint32_t name_len;
  uint32_t cls;
  uint32_t single;
  uint32_t pad;
};

// One data line of the depth file: kept (the reference keeps the row), its
// name (trimmed) in the file, and contigLen.
struct DepthRow {
  uint64_t name_beg;
  uint64_t contig_len;
  uint32_t name_len;
  uint32_t kept;
};

// Depth-parse flags (device word).
enum : uint32_t {
  kDepthUnsupportedNumber = 1u,  // a number rbg_decimal.h does not convert exactly
  kDepthMissing = 2u,            // a contig without a depth row
  kDepthHashCollision = 4u,      // two names with one 64-bit hash (never expected)
};

uint64_t ig_blocks(uint64_t len);
uint64_t ig_groups(uint64_t nblk);
hipError_t launch_fa_summary(const uint8_t* d, uint64_t len, uint32_t* summ, hipStream_t st);
hipError_t launch_fa_carry(const uint32_t* summ, uint64_t nblk, uint32_t* glast, uint32_t* carry, uint32_t* count,
                           hipStream_t st);
hipError_t launch_fa_starts(const uint8_t* d, uint64_t len, const uint32_t* carry, const uint64_t* off, uint64_t R,
                            uint64_t* starts, hipStream_t st);
hipError_t launch_fa_records(const uint8_t* d, uint64_t len, const uint64_t* starts, uint32_t R, uint64_t min_contig,
                             uint64_t min_small, FaRecord* rec, hipStream_t st);
hipError_t launch_fa_copy(const uint8_t* d, const FaRecord* rec, uint32_t R, uint32_t cls, const uint32_t* slot,
                          const uint64_t* off, uint8_t* out, hipStream_t st);
hipError_t launch_fa_names(const uint8_t* d, const FaRecord* rec, uint32_t R, uint32_t cls, const uint32_t* slot,
                           const uint64_t* noff, uint8_t* names, hipStream_t st);
hipError_t launch_nl_count(const uint8_t* d, uint64_t len, uint32_t* count, hipStream_t st);
hipError_t launch_nl_write(const uint8_t* d, uint64_t len, const uint64_t* off, uint64_t* nlpos, hipStream_t st);
hipError_t launch_depth_rows(const uint8_t* d, uint64_t len, const uint64_t* nlpos, uint64_t n_nl, uint64_t n_lines,
                             uint32_t S, uint32_t cv_ext, uint64_t min_depth_len, DepthRow* rows, float* means,
                             uint32_t* flags, hipStream_t st);
hipError_t launch_depth_insert(const uint8_t* d, const DepthRow* rows, uint64_t n_lines, unsigned long long* keys,
                               uint32_t* vals, uint64_t mask, hipStream_t st);
hipError_t launch_depth_lookup(const uint8_t* d, const DepthRow* rows, const float* means, uint32_t S,
                               const unsigned long long* keys, const uint32_t* vals, uint64_t mask,
                               const uint8_t* names, const uint64_t* noff, uint32_t n, float* out, uint32_t* flags,
                               uint32_t* first_missing, hipStream_t st);
hipError_t launch_fill_ig(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);

}  // namespace rbg
