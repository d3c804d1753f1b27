// C ABI query entry points of libebi_hip: Decoder::{materialize, filter,
// filter_materialize, sum, min, max} (lemolatoon/ebi ebi/src/api/
// decoder.rs:202-439) over a resident column.
//
// Flow of one call: upload the optional bitmask, select chunks on the device
// (chunk_id and bitmask pruning, api/decoder.rs:216-226), run the in-situ
// kernels (BUFF, DeltaSprintz) or decode into HBM and run the record-tile
// kernels (other codecs), then copy the small result back.  Per-chunk
// partial results are combined on the host in chunk order, as the Decoder
// loop does.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "../../../include/ebi_hip.h"
#include "../kernels/launch.hpp"
#include "column.hpp"
#include "internal.hpp"

namespace ebi {
namespace {

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

inline int64_t sat_i64(double d) {
  if (std::isnan(d)) return 0;
  if (d >= 9223372036854775808.0) return INT64_MAX;
  if (d < -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}
// Rust f64::min / f64::max: a NaN operand yields the other operand.
double rmin(double a, double b) { return std::isnan(a) ? b : std::isnan(b) ? a : (b < a ? b : a); }
double rmax(double a, double b) { return std::isnan(a) ? b : std::isnan(b) ? a : (b > a ? b : a); }

struct Scratch {
  uint64_t words = 0;  // ceil(N / 64)
  uint64_t *mask = nullptr;
  uint64_t *bits = nullptr;
  uint32_t *sel = nullptr;
  ChunkPred *cpred = nullptr;
  uint64_t *acc = nullptr;
  double *partials = nullptr;
  double *chunk_sum = nullptr;
  uint64_t *block = nullptr;
  double *values = nullptr;
  double *out = nullptr;
};

int prepare(ebi_hip_ctx *ctx, Column &c, const uint64_t *bitset, int64_t chunk_id, bool values, bool out,
            Scratch &s) {
  int rc = check_chunk_errors(ctx, c, chunk_id);
  if (rc) return rc;
  const uint64_t N = c.meta.number_of_records;
  const uint64_t nc = c.chunks.size();
  s.words = (N + 63) / 64;
  const uint64_t wbytes = round_up((s.words + 2) * 8, kAlign);
  const uint64_t nb = launch::compact_blocks(s.words) + 1;
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = off;
    off += round_up(std::max<uint64_t>(bytes, 8), kAlign);
    return o;
  };
  const uint64_t o_mask = take(wbytes), o_bits = take(wbytes), o_sel = take(nc * 4),
                 o_cpred = take(nc * sizeof(ChunkPred)), o_acc = take(nc * 32),
                 o_part = take(c.n_rtiles * 16), o_csum = take(nc * 16), o_block = take(nb * 8);
  if (ctx->scratch_misc.ensure(off)) return fail(ctx, EBI_OOM, "scratch allocation");
  uint8_t *base = (uint8_t *)ctx->scratch_misc.p;
  s.mask = (uint64_t *)(base + o_mask);
  s.bits = (uint64_t *)(base + o_bits);
  s.sel = (uint32_t *)(base + o_sel);
  s.cpred = (ChunkPred *)(base + o_cpred);
  s.acc = (uint64_t *)(base + o_acc);
  s.partials = (double *)(base + o_part);
  s.chunk_sum = (double *)(base + o_csum);
  s.block = (uint64_t *)(base + o_block);
  if (values) {
    if (ctx->scratch_values.ensure(std::max<uint64_t>(N, 1) * 8)) return fail(ctx, EBI_OOM, "value scratch");
    s.values = (double *)ctx->scratch_values.p;
  }
  if (out) {
    if (ctx->scratch_out.ensure(std::max<uint64_t>(N, 1) * 8)) return fail(ctx, EBI_OOM, "output scratch");
    s.out = (double *)ctx->scratch_out.p;
  }
  if ((rc = upload_view(ctx, c, s.values))) return rc;
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  if (bitset) {
    launch::zero_u32(ctx->stream, (uint32_t *)s.mask, (uint32_t)((s.words + 2) * 2));
    hipError_t e = hipMemcpyAsync(s.mask, bitset, s.words * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "bitmask upload");
  }
  launch::chunk_select(ctx->stream, c.d_view, bitset ? s.mask : nullptr, chunk_id, s.sel, (uint32_t)nc);
  return EBI_OK;
}

QueryArgs args_of(const Scratch &s, bool has_mask, const ebi_pred *p, int op) {
  QueryArgs a{};
  a.mask = has_mask ? s.mask : nullptr;
  a.out_bits = s.bits;
  a.values = s.values;
  a.chunk_sel = s.sel;
  a.cpred = s.cpred;
  a.acc = s.acc;
  a.partials = s.partials;
  a.op = op;
  if (p) {
    a.kind = p->kind;
    a.start_kind = p->start_kind;
    a.end_kind = p->end_kind;
    a.value = p->value;
    a.start = p->start;
    a.end = p->end;
  }
  return a;
}

// Per-chunk BUFF predicate parameters: the early exits of
// buff/filter_cmp.rs:328-351 and buff/filter_eq.rs:103-112, then the
// predicate delta in the chunk's fixed-point domain.
ChunkPred buff_chunk_pred(const DevChunk &d, const ebi_pred &p, bool has_mask) {
  ChunkPred cp{};
  const int32_t frac = (int32_t)d.frac;
  const uint64_t wmask = d.bits >= 64 ? ~0ull : ((1ull << d.bits) - 1);
  if (p.kind == EBI_PRED_EQ || p.kind == EBI_PRED_NE) {
    const int64_t fp = buff_fixed(p.value, frac);
    if (fp < d.base) {
      cp.lmode = kSideEmpty;
    } else if (d.bits == 0) {
      cp.lmode = fp == d.base ? kSideAll : kSideEmpty;
    } else {
      cp.lmode = kSideCmp;
      const uint64_t dp = (uint64_t)fp - (uint64_t)d.base;
      if (d.bits < 8 && !has_mask) {
        cp.eq_u8 = 1;  // filter_eq.rs:188-195 compares with `dp as u8`
        cp.ldp = dp;
      } else {
        cp.ldp = dp & wmask;  // planes and tail see only the low B bits
      }
    }
    return cp;
  }
  auto side = [&](int kind, double v, bool greater, uint32_t *mode, uint64_t *dpo, uint32_t *incl) {
    if (kind == EBI_BOUND_NONE) {
      *mode = kSideNone;
      return;
    }
    const bool inc = kind == EBI_BOUND_INCLUSIVE;
    *incl = inc;
    const int64_t fp = buff_fixed(v, frac);
    if (fp < d.base) {
      *mode = greater ? kSideAll : kSideEmpty;
      return;
    }
    if (d.bits == 0) {
      const bool all = greater ? (inc ? d.base >= fp : d.base > fp) : (inc ? d.base <= fp : d.base < fp);
      *mode = all ? kSideAll : kSideEmpty;
      return;
    }
    const uint64_t dp = (uint64_t)fp - (uint64_t)d.base;
    const uint32_t pbits = dp == 0 ? 0 : 64 - (uint32_t)__builtin_clzll(dp);
    if (d.bits < pbits) {
      *mode = greater ? kSideEmpty : kSideAll;
      return;
    }
    *mode = kSideCmp;
    *dpo = dp;
  };
  side(p.start_kind, p.start, true, &cp.lmode, &cp.ldp, &cp.lincl);
  side(p.end_kind, p.end, false, &cp.rmode, &cp.rdp, &cp.rincl);
  return cp;
}

// DeltaSprintz: predicate.map(|v| (v * scale).round() as i64), sprintz.rs:116.
ChunkPred sprintz_chunk_pred(const DevChunk &d, const ebi_pred &p) {
  ChunkPred cp{};
  const double sc = (double)d.scale;
  if (p.kind == EBI_PRED_EQ || p.kind == EBI_PRED_NE) {
    cp.ldp = (uint64_t)sat_i64(std::round(p.value * sc));
    return cp;
  }
  cp.lmode = (uint32_t)p.start_kind;
  cp.rmode = (uint32_t)p.end_kind;
  cp.ldp = (uint64_t)sat_i64(std::round(p.start * sc));
  cp.rdp = (uint64_t)sat_i64(std::round(p.end * sc));
  return cp;
}

int upload_cpred(ebi_hip_ctx *ctx, Column &c, const ebi_pred &p, bool has_mask, const Scratch &s) {
  const size_t nc = c.chunks.size();
  std::vector<ChunkPred> cps(nc);
  for (size_t k = 0; k < nc; ++k) {
    const DevChunk &d = c.chunks[k].dev;
    cps[k] = c.meta.scheme == kBUFF ? buff_chunk_pred(d, p, has_mask) : sprintz_chunk_pred(d, p);
  }
  if (!nc) return EBI_OK;
  hipError_t e = hipMemcpyAsync(s.cpred, cps.data(), nc * sizeof(ChunkPred), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // cps is a host temporary
  if (e != hipSuccess) return hip_fail(ctx, e, "predicate upload");
  return EBI_OK;
}

bool in_situ(const Column &c) { return c.meta.scheme == kBUFF || c.meta.scheme == kDeltaSprintz; }

int finish(ebi_hip_ctx *ctx, const char *what) {
  hipEventRecord(ctx->ev1, ctx->stream);
  int rc = sync_and_check(ctx, what);
  if (rc) return rc;
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->last_ms = ms;
  return EBI_OK;
}

int copy_back(ebi_hip_ctx *ctx, void *dst, const void *src, size_t bytes, const char *what) {
  if (!bytes) return EBI_OK;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return EBI_OK;
}

// op: 0 sum, 1 min, 2 max
int aggregate(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const uint64_t *bitset, int64_t chunk_id, int op,
              double *out) {
  if (!colc || !out) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  const bool insitu = in_situ(c);
  Scratch s;
  int rc = prepare(ctx, c, bitset, chunk_id, !insitu, false, s);
  if (rc) return rc;
  hipEventRecord(ctx->ev0, ctx->stream);
  if (!insitu && (rc = enqueue_decode(ctx, c))) return rc;
  const uint32_t nc = (uint32_t)c.chunks.size();
  QueryArgs a = args_of(s, bitset != nullptr, nullptr, op);
  launch::aggregate(ctx->stream, c.d_view, c.d_rec_jobs, c.n_rec_jobs, a, nc);
  if (!insitu && op == 0) launch::chunk_dd(ctx->stream, c.d_view, s.sel, s.partials, s.chunk_sum, nc);
  if ((rc = finish(ctx, "aggregate"))) return rc;
  std::vector<uint32_t> sel(nc);
  std::vector<uint64_t> acc(4ull * nc);
  std::vector<double> csum(2ull * nc);
  if ((rc = copy_back(ctx, sel.data(), s.sel, nc * 4, "aggregate result"))) return rc;
  if ((rc = copy_back(ctx, acc.data(), s.acc, acc.size() * 8, "aggregate result"))) return rc;
  if (!insitu && op == 0 && (rc = copy_back(ctx, csum.data(), s.chunk_sum, csum.size() * 8, "aggregate result")))
    return rc;
  // Decoder fold, api/decoder.rs:323-439.
  double result = op == 0 ? 0.0 : op == 1 ? INFINITY : -INFINITY;
  double hi = 0.0, lo = 0.0;  // double-double total for the default codecs
  for (uint32_t k = 0; k < nc; ++k) {
    if (!sel[k]) continue;
    const DevChunk &d = c.chunks[k].dev;
    const uint64_t *ak = acc.data() + 4ull * k;
    double v;
    if (c.meta.scheme == kBUFF) {
      const double scale = std::ldexp(1.0, (int)d.frac);  // 2.0f64.powi(frac)
      if (op == 0) {
        if (d.bits == 0)  // buff/sum.rs:80-84: value * n, bitmask ignored
          v = ((double)d.base / scale) * (double)d.n;
        else
          v = (double)(int64_t)(ak[0] + (uint64_t)d.base * ak[1]) / scale;
      } else if (d.bits == 0 || (op == 1 && !bitset)) {
        v = (double)d.base / scale;  // min_max.rs:316-341 and the B = 0 exits
      } else {
        const uint64_t dd = op == 1 ? ak[2] : ak[3];
        v = (double)(int64_t)((uint64_t)d.base + dd) / scale;
      }
    } else if (c.meta.scheme == kDeltaSprintz) {
      const double sc = (double)d.scale;
      if (bitset && ak[1] == 0 && op == 0) v = NAN;
      else if (op == 0) v = (double)(int64_t)ak[0] / sc;
      else v = (double)(int64_t)((op == 1 ? ak[2] : ak[3]) ^ 0x8000000000000000ull) / sc;
    } else {
      if (op == 0) {
        // keep full precision across chunks (two-sum)
        for (int t = 0; t < 2; ++t) {
          const double b = csum[2ull * k + t];
          const double sum = hi + b;
          if (!std::isfinite(sum)) {
            hi = sum;
            lo = 0.0;
            continue;
          }
          const double bv = sum - hi;
          const double e = (hi - (sum - bv)) + (b - bv);
          hi = sum;
          lo += e;
        }
        continue;
      }
      const uint64_t key = op == 1 ? ak[2] : ak[3];
      if ((op == 1 && key == ~0ull) || (op == 2 && key == 0)) {
        v = op == 1 ? INFINITY : -INFINITY;
      } else {
        const uint64_t b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
        std::memcpy(&v, &b, 8);
      }
    }
    if (op == 0) result += v;
    else if (op == 1) result = rmin(result, v);
    else result = rmax(result, v);
  }
  if (!insitu && op == 0) result = std::isfinite(hi) ? hi + lo : hi;
  *out = result;
  return EBI_OK;
}

}  // namespace
}  // namespace ebi

using namespace ebi;


extern "C" int ebi_hip_sum(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset, int64_t chunk_id,
                           double *out) {
  EBI_GUARD(ctx);
  return aggregate(ctx, col, bitset, chunk_id, 0, out);
}
extern "C" int ebi_hip_min(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset, int64_t chunk_id,
                           double *out) {
  EBI_GUARD(ctx);
  return aggregate(ctx, col, bitset, chunk_id, 1, out);
}
extern "C" int ebi_hip_max(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset, int64_t chunk_id,
                           double *out) {
  EBI_GUARD(ctx);
  return aggregate(ctx, col, bitset, chunk_id, 2, out);
}

extern "C" int ebi_hip_filter(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const ebi_pred *pred, const uint64_t *bitset,
                              int64_t chunk_id, uint64_t *out_bitset) {
  EBI_GUARD(ctx);
  if (!colc || !pred || (!out_bitset && colc->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
  if (pred->kind < 0 || pred->kind > 2) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  const bool insitu = in_situ(c);
  Scratch s;
  int rc = prepare(ctx, c, bitset, chunk_id, !insitu, false, s);
  if (rc) return rc;
  if (insitu && (rc = upload_cpred(ctx, c, *pred, bitset != nullptr, s))) return rc;
  hipEventRecord(ctx->ev0, ctx->stream);
  launch::zero_u32(ctx->stream, (uint32_t *)s.bits, (uint32_t)((s.words + 2) * 2));
  if (!insitu && (rc = enqueue_decode(ctx, c))) return rc;
  launch::filter(ctx->stream, c.d_view, c.d_rec_jobs, c.n_rec_jobs, args_of(s, bitset != nullptr, pred, 0), 1);
  if ((rc = finish(ctx, "filter"))) return rc;
  return copy_back(ctx, out_bitset, s.bits, s.words * 8, "filter result");
}

// Compact the records of `values` whose bit is set in s.bits to the host.
static int compact_to_host(ebi_hip_ctx *ctx, Scratch &s, double *out, size_t *n_out) {
  uint64_t *d_n = s.block + launch::compact_blocks(s.words);
  launch::compact(ctx->stream, s.bits, s.words, s.values, s.out, s.block, d_n);
  int rc = finish(ctx, "materialize");
  if (rc) return rc;
  uint64_t n = 0;
  if ((rc = copy_back(ctx, &n, d_n, 8, "materialize count"))) return rc;
  if (s.words == 0) n = 0;
  if ((rc = copy_back(ctx, out, s.out, n * 8, "materialize values"))) return rc;
  *n_out = (size_t)n;
  return EBI_OK;
}

extern "C" int ebi_hip_materialize(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const uint64_t *bitset,
                                   int64_t chunk_id, double *out, size_t *n_out) {
  EBI_GUARD(ctx);
  if (!colc || !n_out || (!out && colc->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  *n_out = 0;
  Scratch s;
  int rc = prepare(ctx, c, bitset, chunk_id, true, bitset != nullptr, s);
  if (rc) return rc;
  hipEventRecord(ctx->ev0, ctx->stream);
  if ((rc = enqueue_decode(ctx, c))) return rc;
  if (!bitset) {
    // Every record of the selected chunks, in chunk order.
    if ((rc = finish(ctx, "materialize"))) return rc;
    uint64_t lo = 0, hi = c.meta.number_of_records;
    if (chunk_id >= 0) {
      if ((uint64_t)chunk_id >= c.chunks.size()) return EBI_OK;
      lo = c.chunks[chunk_id].dev.rec_off;
      hi = lo + c.chunks[chunk_id].dev.n;
    }
    if ((rc = copy_back(ctx, out, s.values + lo, (hi - lo) * 8, "materialize values"))) return rc;
    *n_out = (size_t)(hi - lo);
    return EBI_OK;
  }
  launch::zero_u32(ctx->stream, (uint32_t *)s.bits, (uint32_t)((s.words + 2) * 2));
  launch::filter(ctx->stream, c.d_view, c.d_rec_jobs, c.n_rec_jobs, args_of(s, true, nullptr, 0), 0);
  return compact_to_host(ctx, s, out, n_out);
}

extern "C" int ebi_hip_filter_materialize(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const ebi_pred *pred,
                                          const uint64_t *bitset, int64_t chunk_id, double *out, size_t *n_out) {
  EBI_GUARD(ctx);
  if (!colc || !pred || !n_out || (!out && colc->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
  if (pred->kind < 0 || pred->kind > 2) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  *n_out = 0;
  Scratch s;
  int rc = prepare(ctx, c, bitset, chunk_id, true, true, s);
  if (rc) return rc;
  // BUFF filters in its fixed-point domain, then materializes the filtered
  // records (chunk_reader/buff.rs:230-241); every other codec, DeltaSprintz
  // included, evaluates the f64 predicate on decoded records
  // (query.rs:64-88).
  const bool buff = c.meta.scheme == kBUFF;
  if (buff && (rc = upload_cpred(ctx, c, *pred, bitset != nullptr, s))) return rc;
  hipEventRecord(ctx->ev0, ctx->stream);
  launch::zero_u32(ctx->stream, (uint32_t *)s.bits, (uint32_t)((s.words + 2) * 2));
  if ((rc = enqueue_decode(ctx, c))) return rc;
  QueryArgs a = args_of(s, bitset != nullptr, pred, 0);
  launch::filter(ctx->stream, c.d_view, c.d_rec_jobs, c.n_rec_jobs, a, buff ? 1 : 2);
  return compact_to_host(ctx, s, out, n_out);
}
