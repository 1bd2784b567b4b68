// librbg: ingest entry points (rbg_ingest_*): the FASTA and depth-file parses
// on the device (kernels in rbg_ingest.hip) and the resident assembly built
// from them.
//
// FASTA: the file is copied to HBM; per 4096-byte block the record starts
// are counted, a look-back and a scan place them, one wave per record finds
// its name, header end and body and counts its kept bases
// (fa_records_kernel).  The records' metadata (48 bytes each) come back to
// the host, which numbers the large and small contigs in file order and lays
// out their offsets; the large contigs' bases are then compacted on the
// device (fa_copy_kernel) and their names gathered there.
//
// Depth file: copied to HBM; newline positions by count + scan + write; one
// thread per line parses the row (depth_rows_kernel); the rows' names go into
// a device hash table keeping the first row of each name; each contig looks
// its name up and copies its row.
#include <cstring>
#include <string>
#include <string_view>
#include <unordered_set>
#include <vector>

#include "../../include/rbg.h"
#include "rbg_build.h"
#include "rbg_ingest.h"
#include "rbg_internal.h"
#include "rbg_kernels.h"

using namespace rbg;

struct rbg_ingest {
  uint64_t records = 0;
  uint32_t n_large = 0, n_small = 0;
  uint64_t min_small = 0;
  DevBuf file;                       // the FASTA bytes
  uint64_t file_len = 0;
  DevBuf seqs;                       // large contigs packed (+64 readable bytes)
  std::vector<uint64_t> offs;        // n_large + 1
  std::string names, small_names;    // back to back
  std::vector<uint64_t> name_off, small_name_off, lens, small_lens;
  DevBuf dnames, dname_off, dsnames, dsname_off;  // device copies for the depth lookup
  uint32_t S = 0;
  uint64_t depth_rows = 0;
  DevBuf depth, small_depth;         // n x S raw means
  double ms_fasta = 0, ms_depth = 0;
};

namespace {

float event_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// Exclusive scan of n u32 counts into n + 1 u64 offsets (device).
rbg_status scan_counts(rbg_ctx* ctx, const uint32_t* cnt, uint64_t* off, uint32_t n) {
  size_t tmp = 0;
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64(nullptr, nullptr, n, nullptr, tmp, ctx->stream));
  DevBuf t;
  RBG_TRY_HIP(ctx, t.ensure(tmp));
  RBG_TRY_HIP(ctx, exclusive_scan_u32_to_u64(cnt, off, n, t.p, tmp, ctx->stream));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RBG_OK;
}

// Device -> host copy on the context's stream (non-blocking: the legacy
// default stream does not wait for it), waited for.
rbg_status d2h(rbg_ctx* ctx, void* dst, const void* src, size_t bytes) {
  RBG_TRY_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RBG_OK;
}

}  // namespace

extern "C" {

rbg_status rbg_ingest_fasta(rbg_ctx* ctx, const char* data, uint64_t len, uint64_t min_contig,
                            uint64_t min_small_contig, rbg_ingest** out) {
  if (!ctx || !out || (len && !data)) return RBG_ERR_ARG;
  *out = nullptr;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  auto ig = std::make_unique<rbg_ingest>();
  ig->min_small = min_small_contig;
  ig->offs.assign(1, 0);
  ig->name_off.assign(1, 0);
  ig->small_name_off.assign(1, 0);
  RBG_TRY_HIP(ctx, hipEventRecord(ctx->ev[0], st));
  if (len == 0) {
    RBG_TRY_HIP(ctx, ig->seqs.ensure(64));
    *out = ig.release();
    return RBG_OK;
  }
  RBG_TRY_HIP(ctx, ig->file.ensure(len + 16));
  ig->file_len = len;
  {
    const H2DPart part{ig->file.p, data, len};
    RBG_TRY(h2d_parts(ctx, &part, 1));
  }
  const uint8_t* d = (const uint8_t*)ig->file.p;
  const uint64_t nblk = ig_blocks(len);
  if (nblk >= (1ull << 32)) return ctx->fail(RBG_ERR_UNSUPPORTED, "FASTA larger than 16 TiB");
  DevBuf summ, glast, carry, count, boff;
  RBG_TRY_HIP(ctx, summ.ensure(nblk * 3 * 4));
  RBG_TRY_HIP(ctx, glast.ensure(ig_groups(nblk) * 4));
  RBG_TRY_HIP(ctx, carry.ensure(nblk * 4));
  RBG_TRY_HIP(ctx, count.ensure(nblk * 4));
  RBG_TRY_HIP(ctx, boff.ensure((nblk + 1) * 8));
  RBG_TRY_HIP(ctx, launch_fa_summary(d, len, (uint32_t*)summ.p, st));
  RBG_TRY_HIP(ctx, launch_fa_carry((const uint32_t*)summ.p, nblk, (uint32_t*)glast.p, (uint32_t*)carry.p,
                                   (uint32_t*)count.p, st));
  RBG_TRY(scan_counts(ctx, (const uint32_t*)count.p, (uint64_t*)boff.p, (uint32_t)nblk));
  uint64_t R = 0;
  RBG_TRY(d2h(ctx, &R, (uint64_t*)boff.p + nblk, 8));
  if (R >= (1ull << 32)) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 2^32 FASTA records");
  DevBuf starts, recs;
  RBG_TRY_HIP(ctx, starts.ensure(std::max<uint64_t>(R, 1) * 8));
  RBG_TRY_HIP(ctx, recs.ensure(std::max<uint64_t>(R, 1) * sizeof(FaRecord)));
  RBG_TRY_HIP(ctx, launch_fa_starts(d, len, (const uint32_t*)carry.p, (const uint64_t*)boff.p, R,
                                    (uint64_t*)starts.p, st));
  RBG_TRY_HIP(ctx, launch_fa_records(d, len, (const uint64_t*)starts.p, (uint32_t)R, min_contig, min_small_contig,
                                     (FaRecord*)recs.p, st));
  std::vector<FaRecord> rh(R);
  if (R) RBG_TRY_HIP(ctx, hipMemcpyAsync(rh.data(), recs.p, R * sizeof(FaRecord), hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  // number the large and small contigs in file order; a header without a
  // newline ends the parse (it can only be the last record)
  std::vector<uint32_t> slot(R, 0), sslot(R, 0);
  uint64_t at = 0, nat = 0, snat = 0;
  for (uint64_t r = 0; r < R; ++r) {
    const FaRecord& x = rh[r];
    if (x.cls == kFaDropped) break;
    ++ig->records;
    if (x.cls == kFaLarge) {
      slot[r] = ig->n_large++;
      at += x.seq_len;
      nat += x.name_len;
      ig->offs.push_back(at);
      ig->name_off.push_back(nat);
      ig->lens.push_back(x.seq_len);
    } else if (x.cls == kFaSmall) {
      sslot[r] = ig->n_small++;
      snat += x.name_len;
      ig->small_name_off.push_back(snat);
      ig->small_lens.push_back(x.seq_len);
    }
  }
  if (ig->n_large >= (1u << 24)) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 2^24 large contigs");
  DevBuf dslot, doffs;
  RBG_TRY_HIP(ctx, dslot.ensure(std::max<uint64_t>(R, 1) * 4));
  RBG_TRY_HIP(ctx, doffs.ensure(ig->offs.size() * 8));
  RBG_TRY_HIP(ctx, ig->seqs.ensure(at + 64));
  if (R) RBG_TRY_HIP(ctx, hipMemcpyAsync(dslot.p, slot.data(), R * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(doffs.p, ig->offs.data(), ig->offs.size() * 8, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, launch_fa_copy(d, (const FaRecord*)recs.p, (uint32_t)R, kFaLarge, (const uint32_t*)dslot.p,
                                  (const uint64_t*)doffs.p, (uint8_t*)ig->seqs.p, st));
  // names, gathered on the device (kept there for the depth lookup)
  RBG_TRY_HIP(ctx, ig->dnames.ensure(std::max<uint64_t>(nat, 1)));
  RBG_TRY_HIP(ctx, ig->dname_off.ensure(ig->name_off.size() * 8));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(ig->dname_off.p, ig->name_off.data(), ig->name_off.size() * 8,
                                  hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, launch_fa_names(d, (const FaRecord*)recs.p, (uint32_t)R, kFaLarge, (const uint32_t*)dslot.p,
                                   (const uint64_t*)ig->dname_off.p, (uint8_t*)ig->dnames.p, st));
  RBG_TRY_HIP(ctx, ig->dsnames.ensure(std::max<uint64_t>(snat, 1)));
  RBG_TRY_HIP(ctx, ig->dsname_off.ensure(ig->small_name_off.size() * 8));
  RBG_TRY_HIP(ctx, hipMemcpyAsync(ig->dsname_off.p, ig->small_name_off.data(), ig->small_name_off.size() * 8,
                                  hipMemcpyHostToDevice, st));
  if (R) RBG_TRY_HIP(ctx, hipMemcpyAsync(dslot.p, sslot.data(), R * 4, hipMemcpyHostToDevice, st));
  RBG_TRY_HIP(ctx, launch_fa_names(d, (const FaRecord*)recs.p, (uint32_t)R, kFaSmall, (const uint32_t*)dslot.p,
                                   (const uint64_t*)ig->dsname_off.p, (uint8_t*)ig->dsnames.p, st));
  ig->names.resize(nat);
  ig->small_names.resize(snat);
  if (nat) RBG_TRY_HIP(ctx, hipMemcpyAsync(&ig->names[0], ig->dnames.p, nat, hipMemcpyDeviceToHost, st));
  if (snat) RBG_TRY_HIP(ctx, hipMemcpyAsync(&ig->small_names[0], ig->dsnames.p, snat, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipEventRecord(ctx->ev[1], st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  ig->ms_fasta = event_ms(ctx->ev[0], ctx->ev[1]);
  ig->file.release();  // the bases are packed; the raw file is no longer needed
  *out = ig.release();
  return RBG_OK;
}

rbg_status rbg_ingest_depth(rbg_ctx* ctx, rbg_ingest* ig, const char* data, uint64_t len, uint32_t S, int cv_ext,
                            uint64_t min_depth_len) {
  if (!ctx || !ig || (len && !data)) return RBG_ERR_ARG;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // duplicate large-contig names: the reference keeps the first and drops the
  // rest of the contig set when a depth file is given; refused here
  {
    std::unordered_set<std::string_view> seen;
    seen.reserve(ig->n_large * 2 + 1);
    for (uint32_t i = 0; i < ig->n_large; ++i) {
      const std::string_view nm(ig->names.data() + ig->name_off[i], ig->name_off[i + 1] - ig->name_off[i]);
      if (!seen.insert(nm).second)
        return ctx->fail(RBG_ERR_UNSUPPORTED, "duplicate contig name " + std::string(nm) +
                                                  ": the reference drops later duplicates; deduplicate the FASTA");
    }
  }
  if (S == 0) {  // ncols(file, 1) of the first data line, less the non-feature columns
    const char* end = data + len;
    const char* nl0 = len ? (const char*)std::memchr(data, '\n', len) : nullptr;
    uint64_t nc = 0;
    if (nl0 && nl0 + 1 < end) {
      const char* p = nl0 + 1;
      const char* nl1 = (const char*)std::memchr(p, '\n', (size_t)(end - p));
      const char* le = nl1 ? nl1 : end;
      for (const char* f = p; f < le;) {
        const char* t = (const char*)std::memchr(f, '\t', (size_t)(le - f));
        ++nc;
        if (!t) break;
        f = t + 1;
      }
    }
    const uint64_t nonfeat = cv_ext ? 1 : 3;
    if (nc <= nonfeat) return ctx->fail(RBG_ERR_ARG, "depth file: no feature columns on its first data line");
    uint64_t s = nc - nonfeat;
    if (!cv_ext) {
      if (s % 2) return ctx->fail(RBG_ERR_ARG, "depth file: odd number of mean / variance columns");
      s /= 2;
    }
    S = (uint32_t)s;
  }
  if (S > 4096) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 4096 depth samples");
  ig->S = S;
  RBG_TRY_HIP(ctx, hipEventRecord(ctx->ev[2], st));
  DevBuf file;
  RBG_TRY_HIP(ctx, file.ensure(len + 16));
  if (len) {
    const H2DPart part{file.p, data, len};
    RBG_TRY(h2d_parts(ctx, &part, 1));
  }
  const uint8_t* d = (const uint8_t*)file.p;
  const uint64_t nblk = ig_blocks(len);
  DevBuf cnt, off;
  RBG_TRY_HIP(ctx, cnt.ensure(std::max<uint64_t>(nblk, 1) * 4));
  RBG_TRY_HIP(ctx, off.ensure((nblk + 1) * 8));
  uint64_t n_nl = 0;
  DevBuf nlpos;
  if (nblk) {
    RBG_TRY_HIP(ctx, launch_nl_count(d, len, (uint32_t*)cnt.p, st));
    RBG_TRY(scan_counts(ctx, (const uint32_t*)cnt.p, (uint64_t*)off.p, (uint32_t)nblk));
    RBG_TRY(d2h(ctx, &n_nl, (uint64_t*)off.p + nblk, 8));
  }
  RBG_TRY_HIP(ctx, nlpos.ensure(std::max<uint64_t>(n_nl, 1) * 8));
  if (n_nl) RBG_TRY_HIP(ctx, launch_nl_write(d, len, (const uint64_t*)off.p, (uint64_t*)nlpos.p, st));
  // data lines start after each newline but a final one
  uint64_t n_lines = 0;
  if (n_nl) {
    uint64_t last = 0;
    RBG_TRY(d2h(ctx, &last, (uint64_t*)nlpos.p + n_nl - 1, 8));
    n_lines = last + 1 < len ? n_nl : n_nl - 1;
  }
  if (n_lines >= 0xFFFFFFFFull) return ctx->fail(RBG_ERR_UNSUPPORTED, "more than 2^32 depth rows");
  DevBuf rows, means, flags;
  RBG_TRY_HIP(ctx, rows.ensure(std::max<uint64_t>(n_lines, 1) * sizeof(DepthRow)));
  RBG_TRY_HIP(ctx, means.ensure(std::max<uint64_t>(n_lines * S, 1) * 4));
  RBG_TRY_HIP(ctx, flags.ensure(8));
  RBG_TRY_HIP(ctx, launch_fill_ig((uint32_t*)flags.p, 1, 0u, st));
  RBG_TRY_HIP(ctx, launch_fill_ig((uint32_t*)flags.p + 1, 1, 0xFFFFFFFFu, st));
  RBG_TRY_HIP(ctx, launch_depth_rows(d, len, (const uint64_t*)nlpos.p, n_nl, n_lines, S, cv_ext ? 1u : 0u,
                                     min_depth_len, (DepthRow*)rows.p, (float*)means.p, (uint32_t*)flags.p, st));
  uint64_t cap = 1024;
  while (cap < 2 * n_lines) cap <<= 1;
  DevBuf keys, vals;
  RBG_TRY_HIP(ctx, keys.ensure(cap * 8));
  RBG_TRY_HIP(ctx, vals.ensure(cap * 4));
  RBG_TRY_HIP(ctx, launch_fill_ig((uint32_t*)keys.p, cap * 2, 0u, st));
  RBG_TRY_HIP(ctx, launch_fill_ig((uint32_t*)vals.p, cap, 0xFFFFFFFFu, st));
  RBG_TRY_HIP(ctx, launch_depth_insert(d, (const DepthRow*)rows.p, n_lines, (unsigned long long*)keys.p,
                                       (uint32_t*)vals.p, cap - 1, st));
  RBG_TRY_HIP(ctx, ig->depth.ensure(std::max<uint64_t>((uint64_t)ig->n_large * S, 1) * 4));
  RBG_TRY_HIP(ctx, ig->small_depth.ensure(std::max<uint64_t>((uint64_t)ig->n_small * S, 1) * 4));
  RBG_TRY_HIP(ctx, launch_depth_lookup(d, (const DepthRow*)rows.p, (const float*)means.p, S,
                                       (const unsigned long long*)keys.p, (const uint32_t*)vals.p, cap - 1,
                                       (const uint8_t*)ig->dnames.p, (const uint64_t*)ig->dname_off.p, ig->n_large,
                                       (float*)ig->depth.p, (uint32_t*)flags.p, (uint32_t*)flags.p + 1, st));
  uint32_t fl[2] = {0, 0};
  RBG_TRY_HIP(ctx, hipMemcpyAsync(fl, flags.p, 8, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  const uint32_t missing_large = (fl[0] & kDepthMissing) ? fl[1] : 0xFFFFFFFFu;
  // small contigs: a missing row leaves zeros (their rows serve recruitment only)
  RBG_TRY_HIP(ctx, launch_fill_ig((uint32_t*)flags.p + 1, 1, 0xFFFFFFFFu, st));
  RBG_TRY_HIP(ctx, launch_depth_lookup(d, (const DepthRow*)rows.p, (const float*)means.p, S,
                                       (const unsigned long long*)keys.p, (const uint32_t*)vals.p, cap - 1,
                                       (const uint8_t*)ig->dsnames.p, (const uint64_t*)ig->dsname_off.p, ig->n_small,
                                       (float*)ig->small_depth.p, (uint32_t*)flags.p, (uint32_t*)flags.p + 1, st));
  uint32_t fl2[2] = {0, 0};
  RBG_TRY_HIP(ctx, hipMemcpyAsync(fl2, flags.p, 8, hipMemcpyDeviceToHost, st));
  // rows in the map: the slots in use
  std::vector<uint32_t> vh(cap);
  RBG_TRY_HIP(ctx, hipMemcpyAsync(vh.data(), vals.p, cap * 4, hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipEventRecord(ctx->ev[3], st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  ig->ms_depth = event_ms(ctx->ev[2], ctx->ev[3]);
  ig->depth_rows = 0;
  for (uint32_t v : vh) ig->depth_rows += v != 0xFFFFFFFFu;
  if ((fl[0] | fl2[0]) & kDepthUnsupportedNumber)
    return ctx->fail(RBG_ERR_UNSUPPORTED, "depth file: a number outside the exact conversion (hexadecimal, NaN "
                                          "payload, more than 19 significant digits or beyond 1e-30 .. 1e38)");
  if ((fl[0] | fl2[0]) & kDepthHashCollision)
    return ctx->fail(RBG_ERR_UNSUPPORTED, "depth file: two names share a 64-bit hash");
  if (missing_large != 0xFFFFFFFFu) {
    const uint32_t i = missing_large;
    return ctx->fail(RBG_ERR_ARG, "contig " + std::string(ig->names.data() + ig->name_off[i],
                                                          ig->name_off[i + 1] - ig->name_off[i]) +
                                      " has no row in the depth file");
  }
  return RBG_OK;
}

rbg_status rbg_ingest_get(const rbg_ingest* ig, rbg_ingest_info* info) {
  if (!ig || !info) return RBG_ERR_ARG;
  info->records = ig->records;
  info->n_large = ig->n_large;
  info->n_small = ig->n_small;
  info->large_bases = ig->offs.back();
  info->names = ig->names.data();
  info->name_off = ig->name_off.data();
  info->lens = ig->lens.data();
  info->small_names = ig->small_names.data();
  info->small_name_off = ig->small_name_off.data();
  info->small_lens = ig->small_lens.data();
  info->S = ig->S;
  info->depth_rows = ig->depth_rows;
  info->ms_fasta = ig->ms_fasta;
  info->ms_depth = ig->ms_depth;
  return RBG_OK;
}

rbg_status rbg_ingest_copy(rbg_ctx* ctx, const rbg_ingest* ig, char* seqs, uint64_t* offsets, float* depth,
                           float* small_depth) {
  if (!ctx || !ig) return RBG_ERR_ARG;
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  if (seqs && ig->offs.back())
    RBG_TRY_HIP(ctx, hipMemcpyAsync(seqs, ig->seqs.p, ig->offs.back(), hipMemcpyDeviceToHost, st));
  if (offsets) std::memcpy(offsets, ig->offs.data(), ig->offs.size() * 8);
  if (depth && ig->S && ig->n_large)
    RBG_TRY_HIP(ctx, hipMemcpyAsync(depth, ig->depth.p, (size_t)ig->n_large * ig->S * 4, hipMemcpyDeviceToHost, st));
  if (small_depth && ig->S && ig->n_small)
    RBG_TRY_HIP(ctx, hipMemcpyAsync(small_depth, ig->small_depth.p, (size_t)ig->n_small * ig->S * 4,
                                    hipMemcpyDeviceToHost, st));
  RBG_TRY_HIP(ctx, hipStreamSynchronize(st));
  return RBG_OK;
}

rbg_status rbg_ingest_assembly(rbg_ctx* ctx, const rbg_ingest* ig, double min_cv, rbg_asm** out) {
  if (!ctx || !ig || !out) return RBG_ERR_ARG;
  *out = nullptr;
  if (ig->n_large == 0) return ctx->fail(RBG_ERR_ARG, "no large contigs");
  RBG_TRY_HIP(ctx, hipSetDevice(ctx->device));
  // the depth rows feed the host-side row preparation (depth features, pair
  // order, int8 rows); the bases stay on the device
  std::vector<float> depth((size_t)ig->n_large * ig->S);
  if (ig->S) RBG_TRY(d2h(ctx, depth.data(), ig->depth.p, depth.size() * 4));
  AsmInput in;
  in.in = rbg_assembly_in{(const char*)ig->seqs.p, ig->offs.data(), ig->n_large, ig->S ? depth.data() : nullptr,
                          ig->S};
  in.seqs_on_device = true;
  auto b = std::make_unique<rbg_asm>();
  RBG_TRY(batch_upload(ctx, b.get(), &in, 1, min_cv));
  *out = b.release();
  return RBG_OK;
}

void rbg_ingest_free(rbg_ingest* ig) { delete ig; }

}  // extern "C"
