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

// Rust f64::min / f64::max: a NaN operand yields the other operand.
double rmin(double a, double b) { return std::isnan(a) ? b : std::isnan(b) ? a : (b < a ? b : a); }
double rmax(double a, double b) { return std::isnan(a) ? b : std::isnan(b) ? a : (b > a ? b : a); }

bool is_agg(int op) { return op == kOpSum || op == kOpMin || op == kOpMax || op == kOpFilterSum; }
bool has_pred(int op) { return op == kOpFilter || op == kOpFilterMaterialize || op == kOpFilterSum; }

// Offsets of the device scratch of a plan (all 256-byte aligned; ~0 = none).
struct Layout {
  uint64_t mask, bits, sel, acc, csum, cpred, part, block, values, out, total;
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
  const bool agg = ((p.flags >> kPfAggShift) & 3u) != 0;
  L.mask = p.has_mask ? take(wbytes) : ~0ull;
  L.bits = (p.flags & kPfBits) ? take(wbytes) : ~0ull;
  L.sel = take((uint64_t)p.nc * 4);
  L.acc = agg ? take((uint64_t)p.nc * 32) : ~0ull;
  L.csum = (p.flags & kPfFold) ? take((uint64_t)p.nc * 16) : ~0ull;
  L.cpred = p.cpred ? take((uint64_t)p.nc * sizeof(ChunkPred)) : ~0ull;  // written by the selection pass
  L.part = (p.flags & kPfF64Sum) ? take((uint64_t)p.c->n_rtiles * 16) : ~0ull;
  L.block = p.compact ? take(((uint64_t)p.n_blocks + 1) * 8) : ~0ull;
  // DECODE writes the caller's buffer; fused plans keep the records in registers
  L.values = p.decode && p.q.op != kOpDecode ? take(N * 8) : ~0ull;
  L.out = p.compact ? take(N * 8) : ~0ull;
  L.total = off;
  return L;
}

// Pinned staging: the bitmask copy, the FILTER result.
struct PinLayout {
  uint64_t mask, bits, total;
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
  L.total = off;
  return L;
}

template <class T>
T *at(uint8_t *base, uint64_t off) {
  return off == ~0ull ? nullptr : reinterpret_cast<T *>(base + off);
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
  p.compact = (q.op == kOpMaterialize && (p.has_mask || (p.c->truncated && q.chunk_id < 0))) ||
              q.op == kOpFilterMaterialize;
  p.words = (N + 63) / 64;
  p.nc = (uint32_t)p.c->chunks.size();
  p.n_blocks = p.compact ? launch::compact_blocks(p.words) : 0;
  // Gorilla, Chimp, Elf, ElfOnChimp and Chimp128 filters and aggregates run
  // fused: the segment decode feeds the query in registers
  const bool seg_codec = scheme == kGorilla || scheme == kChimp || scheme == kElf || scheme == kElfOnChimp ||
                         scheme == kChimp128;
  const bool fused = seg_codec && p.c->n_seg_jobs;
  // Uncompressed chunks are read in place and RLE is queried per run: no
  // decode for filters and aggregates
  const bool direct = scheme == kUncompressed || (scheme == kRLE && p.c->rle_direct);
  uint32_t f = p.has_mask ? kPfMask : 0u;
  const uint32_t agg = q.op == kOpSum || q.op == kOpFilterSum ? 1u : q.op == kOpMin ? 2u : q.op == kOpMax ? 3u : 0u;
  f |= agg << kPfAggShift;
  switch (q.op) {
    case kOpMaterialize:
      p.decode = true;
      p.filter_mode = 0;  // select by the bitmask
      if (p.compact) f |= kPfTiles | kPfBits | kPfCompact;
      break;
    case kOpFilter:
      f |= kPfPred | kPfBits | (p.insitu || !fused ? kPfTiles : kPfSegs);
      p.decode = !p.insitu && !fused && !direct;
      p.cpred = p.insitu;
      p.filter_mode = 1;  // code domain for BUFF / DeltaSprintz, f64 otherwise
      break;
    case kOpFilterMaterialize:
      // BUFF filters in its fixed-point domain, then materializes the
      // filtered records (chunk_reader/buff.rs:230-241); every other codec,
      // DeltaSprintz included, evaluates the f64 predicate on decoded
      // records (query.rs:64-88).
      f |= kPfTiles | kPfPred | kPfBits | kPfCompact;
      p.decode = true;
      p.cpred = scheme == kBUFF;
      p.filter_mode = scheme == kBUFF ? 1 : 2;
      break;
    case kOpFilterSum:
      // the sum over the records the filter selects, in one pass: as
      // sum(Some(filter(..))), chunks without a selected record are pruned
      f |= kPfPred | kPfAggSel | kPfFold | (p.insitu || !fused ? kPfTiles : kPfSegs);
      p.decode = !p.insitu && !fused && !direct;
      p.cpred = p.insitu;
      p.filter_mode = 1;
      break;
    case kOpDecode:
      p.decode = true;
      break;
    default:  // SUM / MIN / MAX
      f |= kPfFold | (p.insitu || !fused ? kPfTiles : kPfSegs);
      p.decode = !p.insitu && !fused && !direct;
      break;
  }
  if (agg == 1 && !p.insitu) f |= kPfF64Sum;
  if (p.cpred) f |= kPfCpred;
  p.flags = f;
  // the record-tile kernel of the plan: on BUFF / DeltaSprintz codes when it
  // filters in their domain or aggregates, on RLE runs when it needs no
  // decode, else on the bitmask / decoded records (Uncompressed: in place)
  const bool code = p.insitu && (((f & kPfPred) && p.filter_mode == 1) || agg);
  p.tile_kind = code ? (scheme == kBUFF ? kSpanBuff : kSpanSprintz)
                     : (scheme == kRLE && direct && !p.decode && (f & kPfTiles)) ? kSpanRle : kSpanValues;
  if (p.tile_kind == kSpanRle) p.flags |= kPfFieldTiles;
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
  p.acc = at<uint64_t>(dev, L.acc);
  p.chunk_sum = at<double>(dev, L.csum);
  p.d_cpred = at<ChunkPred>(dev, L.cpred);
  p.partials = at<double>(dev, L.part);
  p.block = at<uint64_t>(dev, L.block);
  p.values = p.q.op == kOpDecode ? p.q.out_values : at<double>(dev, L.values);
  p.out = at<double>(dev, L.out);
  const PinLayout P = pin_layout_of(p);
  p.h_mask = p.direct ? const_cast<uint64_t *>(p.q.bitset) : at<uint64_t>(pin, P.mask);
  p.h_bits = p.direct ? p.q.out_bitset : at<uint64_t>(pin, P.bits);
}

void plan_stage(Plan &p) {
  if (p.has_mask && p.words && !p.direct) std::memcpy(p.h_mask, p.q.bitset, p.words * 8);
  p.n_out = 0;
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
    p.n_out = hi - lo;
    if (hi == lo) return false;
    *err = hipMemcpyAsync(p.q.out_values, p.values + lo, (hi - lo) * 8, hipMemcpyDeviceToHost, s);
    return true;
  }
  if (p.compact) {
    p.n_out = p.words ? p.h_result->n_out : 0;
    if (!p.n_out) return false;
    *err = hipMemcpyAsync(p.q.out_values, p.out, p.n_out * 8, hipMemcpyDeviceToHost, s);
    return true;
  }
  return false;
}

int plan_finish(Plan &p) {
  if (p.status) return p.status;
  const int op = p.q.op;
  if (op == kOpMaterialize || op == kOpFilterMaterialize) {
    *p.q.n_out = (size_t)p.n_out;
    return EBI_OK;
  }
  if (op == kOpFilter) {
    if (p.words && !p.direct) std::memcpy(p.q.out_bitset, p.h_bits, p.words * 8);
    return EBI_OK;
  }
  if (op == kOpDecode) return EBI_OK;
  *p.q.out_scalar = p.h_result->scalar;  // the Decoder's fold, computed on the device (k_group_finish)
  return EBI_OK;
}

// Group block, each part 256-byte aligned: views[n] | decode tile prefix[n+1] |
// decode segment prefix[n+1] | PlanDesc[n] | item prefixes (chunks, span
// jobs of each SpanKind, segment jobs, compaction blocks)[kNumPre][n+1] |
// fills[3n] | results[n] | flags[4].
namespace {
enum GroupPrefix : int { kPreChunk = 0, kPreSpan = 1, kPreSeg = 1 + kSpanKinds, kPreBlock = 2 + kSpanKinds, kNumPre = 3 + kSpanKinds };
struct GroupLayout {
  uint64_t views, dpre, dseg, descs, ipre, fills, results, flags, total;
};
GroupLayout group_layout(size_t n) {
  GroupLayout L{};
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = off;
    off += round_up(std::max<uint64_t>(bytes, 8), kAlign);
    return o;
  };
  L.views = take(n * sizeof(ColView));
  L.dpre = take((n + 1) * 4);
  L.dseg = take((n + 1) * 4);
  L.descs = take(n * sizeof(PlanDesc));
  L.ipre = take((uint64_t)kNumPre * (n + 1) * 4);
  L.fills = take(3 * n * sizeof(GroupFill));
  // results and flags back to back: one copy brings both to the host
  L.results = off;
  off += n * sizeof(PlanResult);
  L.flags = off;
  off += 16;
  L.total = round_up(off, kAlign);
  return L;
}
}  // namespace

uint64_t group_block_bytes(size_t n) { return group_layout(n).total; }

void group_bind(PlanGroup &g, uint8_t *dev, uint8_t *pin, size_t cap) {
  g.dev = dev;
  g.pin = pin;
  g.cap = cap;
  const GroupLayout L = group_layout(cap);
  g.d_flags = reinterpret_cast<uint32_t *>(dev + L.flags);
  g.h_flags = reinterpret_cast<uint32_t *>(pin + L.flags);
  for (size_t i = 0; i < g.plans.size(); ++i) {
    Plan &p = *g.plans[i];
    p.d_view = reinterpret_cast<const ColView *>(dev + L.views) + i;
    p.d_result = reinterpret_cast<PlanResult *>(dev + L.results) + i;
    p.h_result = reinterpret_cast<PlanResult *>(pin + L.results) + i;
  }
}

hipError_t group_enqueue(PlanGroup &g, hipStream_t s, int cu_count, uint32_t blocks_per_cu) {
  const size_t n = g.plans.size();
  if (!n) return hipSuccess;
  const GroupLayout L = group_layout(g.cap);
  ColView *views = reinterpret_cast<ColView *>(g.pin + L.views);
  uint32_t *dpre = reinterpret_cast<uint32_t *>(g.pin + L.dpre);
  uint32_t *dseg = reinterpret_cast<uint32_t *>(g.pin + L.dseg);
  PlanDesc *descs = reinterpret_cast<PlanDesc *>(g.pin + L.descs);
  uint32_t *ipre = reinterpret_cast<uint32_t *>(g.pin + L.ipre);
  auto pre = [&](int k) { return ipre + (size_t)k * (n + 1); };
  uint32_t *pchunk = pre(kPreChunk), *pseg = pre(kPreSeg), *pblock = pre(kPreBlock);
  GroupFill *fills = reinterpret_cast<GroupFill *>(g.pin + L.fills);
  PlanResult *results = reinterpret_cast<PlanResult *>(g.pin + L.results);
  uint64_t jobs = 0, segs = 0, chunks = 0, qsegs = 0, blocks = 0, max_words = 0;
  uint64_t tiles[kSpanKinds] = {};  // span jobs per SpanKind
  uint32_t n_fills = 0;
  uint32_t qseg_codecs = 0;  // bit c: segment jobs of codec c to query
  uint32_t dec_codecs = 0;   // bit c: a column of codec c to decode
  bool any_fold = false;
  auto fill = [&](uint64_t *ptr, uint64_t words, uint64_t v0, uint64_t v1, uint64_t v2, uint64_t v3,
                  uint32_t period) {
    if (!ptr || !words) return;
    GroupFill &f = fills[n_fills++];
    f = GroupFill{};
    f.p = ptr;
    f.n = words;
    f.v[0] = v0;
    f.v[1] = v1;
    f.v[2] = v2;
    f.v[3] = v3;
    f.period = period;
    max_words = std::max(max_words, words);
  };
  for (size_t i = 0; i < n; ++i) {
    Plan &p = *g.plans[i];
    const Column &c = *p.c;
    views[i] = view_of(c, p.values);
    dpre[i] = (uint32_t)jobs;
    dseg[i] = (uint32_t)segs;
    if (p.decode) {
      jobs += c.n_tile_jobs;
      segs += c.n_seg_jobs;
      dec_codecs |= 1u << c.meta.scheme;
    }
    PlanDesc &d = descs[i];
    d = PlanDesc{};
    d.a = args_of(p);
    d.view = p.d_view;
    d.span_jobs = p.tile_kind == kSpanRle ? c.d_field_spans : c.d_span_jobs;
    d.chunk_sum = p.chunk_sum;
    d.out = p.out;
    d.block = p.block;
    d.result = p.d_result;
    d.words = p.words;
    d.chunk_id = p.q.chunk_id;
    d.n_chunks = p.nc;
    d.n_span_jobs = p.tile_kind == kSpanRle ? c.n_field_spans : c.n_span_jobs;
    d.n_seg_jobs = c.n_seg_jobs;
    d.n_blocks = p.n_blocks;
    d.flags = p.flags;
    d.filter_mode = p.filter_mode;
    d.scheme = (uint32_t)c.meta.scheme;
    pchunk[i] = (uint32_t)chunks;
    for (int k = 0; k < kSpanKinds; ++k) pre(kPreSpan + k)[i] = (uint32_t)tiles[k];
    pseg[i] = (uint32_t)qsegs;
    pblock[i] = (uint32_t)blocks;
    const bool runs = p.q.op != kOpDecode && !(p.q.op == kOpMaterialize && !p.compact);
    if (runs) chunks += p.nc;
    if (p.flags & kPfTiles) tiles[p.tile_kind] += p.tile_kind == kSpanRle ? c.n_field_spans : c.n_span_jobs;
    if (p.flags & kPfSegs) {
      qsegs += c.n_seg_jobs;
      if (c.n_seg_jobs) qseg_codecs |= 1u << c.meta.scheme;
    }
    if (p.flags & kPfCompact) blocks += p.n_blocks;
    if (p.flags & kPfFold) any_fold = true;
    results[i] = PlanResult{0.0, 0};
    // scratch that must start cleared: the bitset, the bitmask's two spare
    // words (after the copied ones), the per-chunk accumulators
    if (p.flags & kPfBits) fill(p.bits, p.words + 2, 0, 0, 0, 0, 1);
    if (p.has_mask) fill(p.mask + p.words, 2, 0, 0, 0, 0, 1);
    if (p.acc) fill(p.acc, 4ull * p.nc, 0, 0, ~0ull, 0, 4);
  }
  dpre[n] = (uint32_t)jobs;
  dseg[n] = (uint32_t)segs;
  pchunk[n] = (uint32_t)chunks;
  for (int k = 0; k < kSpanKinds; ++k) pre(kPreSpan + k)[n] = (uint32_t)tiles[k];
  pseg[n] = (uint32_t)qsegs;
  pblock[n] = (uint32_t)blocks;
  if (jobs >= (1ull << 32) || segs >= (1ull << 32) || chunks >= (1ull << 32) || tiles[0] + tiles[1] + tiles[2] + tiles[3] >= (1ull << 32) ||
      qsegs >= (1ull << 32) || blocks >= (1ull << 31))
    return hipErrorInvalidValue;
  g.h_flags[0] = g.h_flags[1] = 0;
  hipError_t e = hipMemcpyAsync(g.dev, g.pin, L.total, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  for (Plan *pp : g.plans) {
    Plan &p = *pp;
    if (p.has_mask && p.words &&
        (e = hipMemcpyAsync(p.mask, p.h_mask, p.words * 8, hipMemcpyHostToDevice, s)) != hipSuccess)
      return e;
  }
  const PlanDesc *dd = reinterpret_cast<const PlanDesc *>(g.dev + L.descs);
  const uint32_t *di = reinterpret_cast<const uint32_t *>(g.dev + L.ipre);
  const uint32_t np = (uint32_t)n;
  const uint32_t max_blocks = (uint32_t)cu_count * blocks_per_cu;
  launch::group_prep(s, reinterpret_cast<const GroupFill *>(g.dev + L.fills), n_fills, max_words, dd, di, np,
                     (uint32_t)chunks);
  launch::decode(s, reinterpret_cast<const ColView *>(g.dev + L.views), np,
                 reinterpret_cast<const uint32_t *>(g.dev + L.dpre), (uint32_t)jobs,
                 reinterpret_cast<const uint32_t *>(g.dev + L.dseg), (uint32_t)segs, g.d_flags, max_blocks, nullptr,
                 false, dec_codecs);
  for (int k = 0; k < kSpanKinds; ++k) launch::group_tiles(s, dd, di + (size_t)(kPreSpan + k) * (n + 1), np, (uint32_t)tiles[k], k);
  launch::query_seg(s, dd, di + (size_t)kPreSeg * (n + 1), np, (uint32_t)qsegs, g.d_flags, max_blocks, qseg_codecs);
  launch::group_fold(s, dd, di, np, (uint32_t)chunks, any_fold);
  launch::group_compact(s, dd, di + (size_t)kPreBlock * (n + 1), np, (uint32_t)blocks);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (Plan *pp : g.plans)
    if (pp->q.op == kOpFilter && pp->words &&
        (e = hipMemcpyAsync(pp->h_bits, pp->bits, pp->words * 8, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return e;
  // every plan's result record and the error flags in one copy
  return hipMemcpyAsync(g.pin + L.results, g.dev + L.results, L.flags + 16 - L.results, hipMemcpyDeviceToHost, s);
}

}  // namespace ebi
