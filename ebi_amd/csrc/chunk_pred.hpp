// Per-chunk query parameters of the in-situ codecs, computed on the device
// by the group's selection pass from the chunk header and the predicate
// (one thread per chunk).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "device.hpp"
#include "host/format.hpp"

namespace ebi {

// Rust `as i64` on f64: saturating, NaN -> 0.
__host__ __device__ inline int64_t f64_as_i64(double d) {
  if (d != d) return 0;
  if (d >= 9223372036854775808.0) return INT64_MAX;
  if (d < -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}

// One side of a BUFF range predicate (buff/filter_cmp.rs:328-351): an early
// exit for the whole chunk, or the bound's delta in the chunk's fixed-point
// domain.
__host__ __device__ inline void buff_side(const DevChunk &d, int kind, double v, bool greater, uint32_t *mode,
                                          uint64_t *dpo, uint32_t *incl) {
  if (kind == EBI_BOUND_NONE) {
    *mode = kSideNone;
    return;
  }
  const bool inc = kind == EBI_BOUND_INCLUSIVE;
  *incl = inc;
  const int64_t fp = buff_fixed(v, (int32_t)d.frac);
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
}

// BUFF: the early exits of buff/filter_cmp.rs:328-351 and
// buff/filter_eq.rs:103-112, then the predicate delta.
__host__ __device__ inline ChunkPred buff_chunk_pred(const DevChunk &d, const ebi_pred &p, bool has_mask) {
  ChunkPred cp{};
  const uint64_t wmask = d.bits >= 64 ? ~0ull : ((1ull << d.bits) - 1);
  if (p.kind == EBI_PRED_EQ || p.kind == EBI_PRED_NE) {
    const int64_t fp = buff_fixed(p.value, (int32_t)d.frac);
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
  buff_side(d, p.start_kind, p.start, true, &cp.lmode, &cp.ldp, &cp.lincl);
  buff_side(d, p.end_kind, p.end, false, &cp.rmode, &cp.rdp, &cp.rincl);
  return cp;
}

// DeltaSprintz: predicate.map(|v| (v * scale).round() as i64), sprintz.rs:116.
__host__ __device__ inline ChunkPred sprintz_chunk_pred(const DevChunk &d, const ebi_pred &p) {
  ChunkPred cp{};
  const double sc = (double)d.scale;
  if (p.kind == EBI_PRED_EQ || p.kind == EBI_PRED_NE) {
    cp.ldp = (uint64_t)f64_as_i64(round(p.value * sc));
    return cp;
  }
  cp.lmode = (uint32_t)p.start_kind;
  cp.rmode = (uint32_t)p.end_kind;
  cp.ldp = (uint64_t)f64_as_i64(round(p.start * sc));
  cp.rdp = (uint64_t)f64_as_i64(round(p.end * sc));
  return cp;
}

}  // namespace ebi
