// Query plans shared by the synchronous query entry points and the serving
// scheduler.  Semantics follow lemolatoon/ebi's Decoder (ebi/src/api/
// decoder.rs:202-439) and the codec query executors cited below.
#include "plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../kernels/launch.hpp"
#include "internal.hpp"

namespace ebi {
namespace {

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

int64_t sat_i64(double d) {
  if (std::isnan(d)) return 0;
  if (d >= 9223372036854775808.0) return INT64_MAX;
  if (d < -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}
// Rust f64::min / f64::max: a NaN operand yields the other operand.
double rmin(double a, double b) { return std::isnan(a) ? b : std::isnan(b) ? a : (b < a ? b : a); }
double rmax(double a, double b) { return std::isnan(a) ? b : std::isnan(b) ? a : (b > a ? b : a); }

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

bool is_agg(int op) { return op == kOpSum || op == kOpMin || op == kOpMax || op == kOpFilterSum; }
bool has_pred(int op) { return op == kOpFilter || op == kOpFilterMaterialize || op == kOpFilterSum; }
// ops whose filter writes the bitset `bits`
bool writes_bits(const Plan &p) {
  return p.q.op == kOpFilter || p.q.op == kOpFilterMaterialize || p.q.op == kOpFilterSum ||
         (p.q.op == kOpMaterialize && p.compact);
}

// Offsets of the device scratch of a plan (all 256-byte aligned).
struct Layout {
  uint64_t mask, bits, sel, cpred, acc, part, csum, block, values, out, total;
};

Layout layout_of(const Plan &p) {
  Layout L{};
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = off;
    off += round_up(std::max<uint64_t>(bytes, 8), kAlign);
    return o;
  };
  const uint64_t N = p.c->meta.number_of_records;
  const uint64_t wbytes = (p.words + 2) * 8;
  const uint64_t nb = launch::compact_blocks(p.words) + 1;
  L.mask = take(wbytes);
  L.bits = take(wbytes);
  // the per-chunk results (sel, acc, csum) back to back, laid out as in the
  // pinned staging, so that one copy brings them all to the host
  L.sel = take((uint64_t)p.nc * 4);
  L.acc = take((uint64_t)p.nc * 32);
  L.csum = take((uint64_t)p.nc * 16);
  L.cpred = take((uint64_t)p.nc * sizeof(ChunkPred));
  L.part = take((uint64_t)p.c->n_rtiles * 16);
  L.block = take(nb * 8);
  // DECODE writes the caller's buffer; fused plans keep the records in registers
  L.values = p.decode && !p.fused && p.q.op != kOpDecode ? take(N * 8) : ~0ull;
  L.out = (p.q.op == kOpFilterMaterialize || (p.q.op == kOpMaterialize && p.compact)) ? take(N * 8) : ~0ull;
  L.total = off;
  return L;
}

// Pinned staging: mask copy, bits result, sel, acc, csum, count, cpred.
struct PinLayout {
  uint64_t mask, bits, sel, acc, csum, n, cpred, total;
};

PinLayout pin_layout_of(const Plan &p) {
  PinLayout L{};
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = off;
    off += round_up(std::max<uint64_t>(bytes, 8), 64);
    return o;
  };
  L.mask = p.has_mask && !p.direct ? take(p.words * 8) : ~0ull;
  L.bits = p.q.op == kOpFilter && !p.direct ? take(p.words * 8) : ~0ull;
  // the per-chunk results at the device layout's spacing (kAlign)
  off = round_up(off, kAlign);
  L.sel = off;
  off += round_up(std::max<uint64_t>((uint64_t)p.nc * 4, 8), kAlign);
  L.acc = off;
  off += round_up(std::max<uint64_t>((uint64_t)p.nc * 32, 8), kAlign);
  L.csum = off;
  off += round_up(std::max<uint64_t>((uint64_t)p.nc * 16, 8), kAlign);
  L.n = take(8);
  L.cpred = p.cpred ? take((uint64_t)p.nc * sizeof(ChunkPred)) : ~0ull;
  L.total = off;
  return L;
}

QueryArgs args_of(const Plan &p) {
  QueryArgs a{};
  a.mask = p.has_mask ? p.mask : nullptr;
  a.out_bits = p.bits;
  a.values = p.values;
  a.chunk_sel = p.sel;
  a.cpred = p.d_cpred;
  a.acc = p.acc;
  a.partials = p.partials;
  a.op = p.q.op == kOpMin ? 1 : p.q.op == kOpMax ? 2 : 0;
  if (has_pred(p.q.op)) {
    a.kind = p.q.pred.kind;
    a.start_kind = p.q.pred.start_kind;
    a.end_kind = p.q.pred.end_kind;
    a.value = p.q.pred.value;
    a.start = p.q.pred.start;
    a.end = p.q.pred.end;
  }
  return a;
}

template <class T>
T *at(uint8_t *base, uint64_t off) {
  return off == ~0ull ? nullptr : reinterpret_cast<T *>(base + off);
}

// Async part before the decode: inputs and chunk selection.
hipError_t plan_enqueue_pre(Plan &p, hipStream_t s) {
  if (p.has_mask) {
    launch::zero_u32(s, (uint32_t *)p.mask, (uint32_t)((p.words + 2) * 2));
    hipError_t e = hipMemcpyAsync(p.mask, p.h_mask, p.words * 8, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
  }
  if (p.cpred && p.nc) {
    hipError_t e = hipMemcpyAsync(p.d_cpred, p.h_cpred, (size_t)p.nc * sizeof(ChunkPred), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
  }
  if (writes_bits(p)) launch::zero_u32(s, (uint32_t *)p.bits, (uint32_t)((p.words + 2) * 2));
  launch::chunk_select(s, p.d_view, p.has_mask ? p.mask : nullptr, p.q.chunk_id, p.sel, p.nc);
  return hipGetLastError();
}

// The per-chunk results (sel, acc and, with `csum`, the chunk sums) to the
// pinned staging in one copy: both layouts hold them back to back at the
// same offsets.
void copy_results(const Plan &p, hipStream_t s, bool csum) {
  const uint8_t *a = reinterpret_cast<const uint8_t *>(p.sel);
  const uint8_t *b = csum ? reinterpret_cast<const uint8_t *>(p.chunk_sum + 2ull * p.nc)
                          : reinterpret_cast<const uint8_t *>(p.acc + 4ull * p.nc);
  if (p.nc) hipMemcpyAsync(p.h_sel, p.sel, (size_t)(b - a), hipMemcpyDeviceToHost, s);
}

// Async part after the decode: query kernels and small results to pinned.
// err: the group's device error flags; max_blocks: the persistent grid.
hipError_t plan_enqueue_post(Plan &p, hipStream_t s, uint32_t *err, uint32_t max_blocks) {
  Column &c = *p.c;
  const int op = p.q.op;
  if (p.fused) {
    // decode + filter / aggregate in one pass; afterwards the same chunk
    // selection, per-chunk fold and copies as below
    QueryArgs a = args_of(p);
    const int agg = op == kOpSum || op == kOpFilterSum ? 0 : op == kOpMin ? 1 : op == kOpMax ? 2 : -1;
    launch::query_seg(s, p.d_view, c.n_seg_jobs, a, has_pred(op), writes_bits(p), agg, p.nc, err, max_blocks);
    if (op == kOpFilter) {
      if (p.words) hipMemcpyAsync(p.h_bits, p.bits, p.words * 8, hipMemcpyDeviceToHost, s);
      return hipGetLastError();
    }
    if (op == kOpFilterSum) launch::chunk_select(s, p.d_view, p.bits, p.q.chunk_id, p.sel, p.nc);
    if (agg == 0) launch::chunk_dd(s, p.d_view, p.sel, p.partials, p.chunk_sum, p.nc);
    copy_results(p, s, agg == 0);
    return hipGetLastError();
  }
  if (is_agg(op)) {
    QueryArgs a = args_of(p);
    if (op == kOpFilterSum) {
      // filter into `bits`, then sum over them exactly as sum(Some(bits))
      // does: chunks re-selected (pruned) by the filter's bitset
      launch::filter(s, p.d_view, c.d_rec_jobs, c.n_rec_jobs, a, p.filter_mode);
      launch::chunk_select(s, p.d_view, p.bits, p.q.chunk_id, p.sel, p.nc);
      a.mask = p.bits;
    }
    launch::aggregate(s, p.d_view, c.d_rec_jobs, c.n_rec_jobs, a, p.nc);
    const bool f64_sum = !p.insitu && (op == kOpSum || op == kOpFilterSum);
    if (f64_sum) launch::chunk_dd(s, p.d_view, p.sel, p.partials, p.chunk_sum, p.nc);
    copy_results(p, s, f64_sum);
  } else if (op == kOpFilter) {
    launch::filter(s, p.d_view, c.d_rec_jobs, c.n_rec_jobs, args_of(p), p.filter_mode);
    if (p.words) hipMemcpyAsync(p.h_bits, p.bits, p.words * 8, hipMemcpyDeviceToHost, s);
  } else if (op == kOpFilterMaterialize || (op == kOpMaterialize && p.compact)) {
    launch::filter(s, p.d_view, c.d_rec_jobs, c.n_rec_jobs, args_of(p), p.filter_mode);
    launch::compact(s, p.bits, p.words, p.values, p.out, p.block, p.d_n);
    hipMemcpyAsync(p.h_n, p.d_n, 8, hipMemcpyDeviceToHost, s);
  }
  return hipGetLastError();
}

}  // namespace

int plan_init(Plan &p, const ebi_query &q) {
  p = Plan();
  p.q = q;
  if (!q.col) return EBI_INVALID_ARGUMENT;
  p.c = &const_cast<ebi_hip_col *>(q.col)->c;
  const uint64_t N = p.c->meta.number_of_records;
  switch (q.op) {
    case kOpMaterialize:
    case kOpFilterMaterialize:
      if (!q.n_out || (!q.out_values && N)) return EBI_INVALID_ARGUMENT;
      break;
    case kOpFilter:
      if (!q.out_bitset && N) return EBI_INVALID_ARGUMENT;
      break;
    case kOpSum: case kOpMin: case kOpMax: case kOpFilterSum:
      if (!q.out_scalar) return EBI_INVALID_ARGUMENT;
      break;
    case kOpDecode:
      if (!q.out_values && N) return EBI_INVALID_ARGUMENT;
      break;
    default:
      return EBI_INVALID_ARGUMENT;
  }
  if (has_pred(q.op) && (q.pred.kind < 0 || q.pred.kind > 2)) return EBI_INVALID_ARGUMENT;
  // chunks rejected by the load-time validation fail every query that selects them
  for (size_t k = 0; k < p.c->chunks.size(); ++k) {
    if (q.chunk_id >= 0 && (uint64_t)q.chunk_id != k) continue;
    if (p.c->chunks[k].err) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "chunk %zu failed validation at load (flags 0x%x)", k, p.c->chunks[k].err);
      p.error = buf;
      return map_dev_err(p.c->chunks[k].err);
    }
  }
  const int scheme = p.c->meta.scheme;
  p.insitu = scheme == kBUFF || scheme == kDeltaSprintz;
  p.has_mask = q.bitset != nullptr;
  // one chunk's records stay contiguous (its DevChunk n is the shortened count)
  p.compact = p.has_mask || (p.c->truncated && q.chunk_id < 0);
  p.words = (N + 63) / 64;
  p.nc = (uint32_t)p.c->chunks.size();
  switch (q.op) {
    case kOpMaterialize:
      p.decode = true;
      p.filter_mode = 0;  // select by the bitmask
      break;
    case kOpFilter:
      p.decode = !p.insitu;
      p.cpred = p.insitu;
      p.filter_mode = 1;  // code domain for BUFF / DeltaSprintz, f64 otherwise
      break;
    case kOpFilterMaterialize:
      // BUFF filters in its fixed-point domain, then materializes the
      // filtered records (chunk_reader/buff.rs:230-241); every other codec,
      // DeltaSprintz included, evaluates the f64 predicate on decoded
      // records (query.rs:64-88).
      p.decode = true;
      p.cpred = scheme == kBUFF;
      p.filter_mode = scheme == kBUFF ? 1 : 2;
      break;
    case kOpFilterSum:
      p.decode = !p.insitu;
      p.cpred = p.insitu;
      p.filter_mode = 1;  // as kOpFilter
      break;
    case kOpDecode:
      p.decode = true;
      break;
    default:
      p.decode = !p.insitu;
      break;
  }
  // Gorilla, Chimp, Elf and Chimp128 filters and aggregates run fused: the
  // segment decode feeds the query in registers
  const bool seg_codec = scheme == kGorilla || scheme == kChimp || scheme == kElf || scheme == kElfOnChimp ||
                         scheme == kChimp128;
  if (seg_codec && p.c->n_seg_jobs && (q.op == kOpFilter || q.op == kOpFilterSum || is_agg(q.op))) {
    p.fused = true;
    p.decode = false;
  }
  return EBI_OK;
}

PlanNeed plan_need(const Plan &p) {
  PlanNeed n;
  n.dev = layout_of(p).total;
  n.pin = pin_layout_of(p).total;
  return n;
}

void plan_bind(Plan &p, uint8_t *dev, uint8_t *pin) {
  const Layout L = layout_of(p);
  p.mask = at<uint64_t>(dev, L.mask);
  p.bits = at<uint64_t>(dev, L.bits);
  p.sel = at<uint32_t>(dev, L.sel);
  p.d_cpred = at<ChunkPred>(dev, L.cpred);
  p.acc = at<uint64_t>(dev, L.acc);
  p.partials = at<double>(dev, L.part);
  p.chunk_sum = at<double>(dev, L.csum);
  p.block = at<uint64_t>(dev, L.block);
  p.d_n = p.block + launch::compact_blocks(p.words);
  p.values = p.q.op == kOpDecode ? p.q.out_values : at<double>(dev, L.values);
  p.out = at<double>(dev, L.out);
  const PinLayout P = pin_layout_of(p);
  p.h_mask = p.direct ? const_cast<uint64_t *>(p.q.bitset) : at<uint64_t>(pin, P.mask);
  p.h_bits = p.direct ? p.q.out_bitset : at<uint64_t>(pin, P.bits);
  p.h_sel = at<uint32_t>(pin, P.sel);
  p.h_acc = at<uint64_t>(pin, P.acc);
  p.h_csum = at<double>(pin, P.csum);
  p.h_n = at<uint64_t>(pin, P.n);
  p.h_cpred = at<ChunkPred>(pin, P.cpred);
}

void plan_stage(Plan &p) {
  if (p.has_mask && p.words && !p.direct) std::memcpy(p.h_mask, p.q.bitset, p.words * 8);
  if (p.cpred) {
    const bool buff = p.c->meta.scheme == kBUFF;
    for (uint32_t k = 0; k < p.nc; ++k) {
      const DevChunk &d = p.c->chunks[k].dev;
      p.h_cpred[k] = buff ? buff_chunk_pred(d, p.q.pred, p.has_mask) : sprintz_chunk_pred(d, p.q.pred);
    }
  }
  *p.h_n = 0;
}

bool plan_enqueue_values(Plan &p, hipStream_t s, hipError_t *err) {
  *err = hipSuccess;
  const int op = p.q.op;
  const Column &c = *p.c;
  if (op == kOpMaterialize && !p.compact) {
    // every record of the selected chunks, in chunk order
    uint64_t lo = 0, hi = c.meta.number_of_records;
    if (p.q.chunk_id >= 0) {
      if ((uint64_t)p.q.chunk_id >= c.chunks.size()) {
        lo = hi = 0;
      } else {
        lo = c.chunks[p.q.chunk_id].dev.rec_off;
        hi = lo + c.chunks[p.q.chunk_id].dev.n;
      }
    }
    *p.h_n = hi - lo;
    if (hi == lo) return false;
    *err = hipMemcpyAsync(p.q.out_values, p.values + lo, (hi - lo) * 8, hipMemcpyDeviceToHost, s);
    return true;
  }
  if (op == kOpFilterMaterialize || op == kOpMaterialize) {
    if (p.words == 0) *p.h_n = 0;
    if (!*p.h_n) return false;
    *err = hipMemcpyAsync(p.q.out_values, p.out, *p.h_n * 8, hipMemcpyDeviceToHost, s);
    return true;
  }
  return false;
}

int plan_finish(Plan &p) {
  if (p.status) return p.status;
  const int op = p.q.op;
  const Column &c = *p.c;
  if (op == kOpMaterialize || op == kOpFilterMaterialize) {
    *p.q.n_out = (size_t)*p.h_n;
    return EBI_OK;
  }
  if (op == kOpFilter) {
    if (p.words && !p.direct) std::memcpy(p.q.out_bitset, p.h_bits, p.words * 8);
    return EBI_OK;
  }
  if (op == kOpDecode) return EBI_OK;
  // Decoder fold over the selected chunks, api/decoder.rs:323-439.
  const int agg = op == kOpMin ? 1 : op == kOpMax ? 2 : 0;
  const bool mask = p.has_mask || op == kOpFilterSum;
  double result = agg == 0 ? 0.0 : agg == 1 ? INFINITY : -INFINITY;
  double hi = 0.0, lo = 0.0;  // double-double total for the default codecs
  for (uint32_t k = 0; k < p.nc; ++k) {
    if (!p.h_sel[k]) continue;
    const DevChunk &d = c.chunks[k].dev;
    const uint64_t *ak = p.h_acc + 4ull * k;
    double v;
    if (c.meta.scheme == kBUFF) {
      const double scale = std::ldexp(1.0, (int)d.frac);  // 2.0f64.powi(frac)
      if (agg == 0) {
        if (d.bits == 0)  // buff/sum.rs:80-84: value * n, bitmask ignored
          v = ((double)d.base / scale) * (double)d.n;
        else
          v = (double)(int64_t)(ak[0] + (uint64_t)d.base * ak[1]) / scale;
      } else if (d.bits == 0 || (agg == 1 && !mask)) {
        v = (double)d.base / scale;  // min_max.rs:316-341 and the B = 0 exits
      } else {
        const uint64_t dd = agg == 1 ? ak[2] : ak[3];
        v = (double)(int64_t)((uint64_t)d.base + dd) / scale;
      }
    } else if (c.meta.scheme == kDeltaSprintz) {
      const double sc = (double)d.scale;
      if (mask && ak[1] == 0 && agg == 0) v = NAN;
      else if (agg == 0) v = (double)(int64_t)ak[0] / sc;
      else v = (double)(int64_t)((agg == 1 ? ak[2] : ak[3]) ^ 0x8000000000000000ull) / sc;
    } else {
      if (agg == 0) {
        // keep full precision across chunks (two-sum)
        for (int t = 0; t < 2; ++t) {
          const double b = p.h_csum[2ull * k + t];
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
      const uint64_t key = agg == 1 ? ak[2] : ak[3];
      if ((agg == 1 && key == ~0ull) || (agg == 2 && key == 0)) {
        v = agg == 1 ? INFINITY : -INFINITY;
      } else {
        const uint64_t b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
        std::memcpy(&v, &b, 8);
      }
    }
    if (agg == 0) result += v;
    else if (agg == 1) result = rmin(result, v);
    else result = rmax(result, v);
  }
  if (!p.insitu && agg == 0) result = std::isfinite(hi) ? hi + lo : hi;
  *p.q.out_scalar = result;
  return EBI_OK;
}

// Group block: views[n] | prefix[n + 1] | prefix_seg[n + 1] | flags[4],
// each part 256-byte aligned.
uint64_t group_block_bytes(size_t n) {
  return round_up(n * sizeof(ColView), kAlign) + 2 * round_up((n + 1) * 4, kAlign) + kAlign;
}

void group_bind(PlanGroup &g, uint8_t *dev, uint8_t *pin) {
  g.dev = dev;
  g.pin = pin;
  const size_t n = g.plans.size();
  const uint64_t flags = round_up(n * sizeof(ColView), kAlign) + 2 * round_up((n + 1) * 4, kAlign);
  g.d_flags = reinterpret_cast<uint32_t *>(dev + flags);
  g.h_flags = reinterpret_cast<uint32_t *>(pin + flags);
  for (size_t i = 0; i < n; ++i) g.plans[i]->d_view = reinterpret_cast<const ColView *>(dev) + i;
}

hipError_t group_enqueue(PlanGroup &g, hipStream_t s, int cu_count, uint32_t blocks_per_cu) {
  const size_t n = g.plans.size();
  if (!n) return hipSuccess;
  const uint64_t vbytes = round_up(n * sizeof(ColView), kAlign);
  const uint64_t pbytes = round_up((n + 1) * 4, kAlign);
  ColView *views = reinterpret_cast<ColView *>(g.pin);
  uint32_t *pre = reinterpret_cast<uint32_t *>(g.pin + vbytes);
  uint32_t *preseg = reinterpret_cast<uint32_t *>(g.pin + vbytes + pbytes);
  uint64_t jobs = 0, segs = 0;
  for (size_t i = 0; i < n; ++i) {
    Plan &p = *g.plans[i];
    views[i] = view_of(*p.c, p.values);
    pre[i] = (uint32_t)jobs;
    preseg[i] = (uint32_t)segs;
    if (p.decode) {
      jobs += p.c->n_tile_jobs;
      segs += p.c->n_seg_jobs;
    }
  }
  pre[n] = (uint32_t)jobs;
  preseg[n] = (uint32_t)segs;
  if (jobs >= (1ull << 32) || segs >= (1ull << 32)) return hipErrorInvalidValue;
  g.h_flags[0] = g.h_flags[1] = 0;
  hipError_t e = hipMemcpyAsync(g.dev, g.pin, vbytes + 2 * pbytes + 8, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  for (Plan *p : g.plans)
    if ((e = plan_enqueue_pre(*p, s)) != hipSuccess) return e;
  const ColView *dv = reinterpret_cast<const ColView *>(g.dev);
  const uint32_t *dpre = reinterpret_cast<const uint32_t *>(g.dev + vbytes);
  launch::decode(s, dv, (uint32_t)n, dpre, (uint32_t)jobs, dpre + pbytes / 4, (uint32_t)segs, g.d_flags,
                 (uint32_t)cu_count * blocks_per_cu);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (Plan *p : g.plans)
    if ((e = plan_enqueue_post(*p, s, g.d_flags, (uint32_t)cu_count * blocks_per_cu)) != hipSuccess) return e;
  return hipMemcpyAsync(g.h_flags, g.d_flags, 4, hipMemcpyDeviceToHost, s);
}

}  // namespace ebi
