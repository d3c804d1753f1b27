// Query kernels of libebi_hip (gfx950): filters to a bitset, aggregates and
// stream compaction.  Included once, by decode.hip (one device TU).
//
// Work unit: a record tile (kTile = 512 records of one chunk, 8 consecutive
// records per lane).  BUFF and DeltaSprintz evaluate predicates and
// aggregates in their code domain (plane deltas, quantized q); the other
// codecs work on records decoded into HBM.  Semantics follow the reference
// readers cited on each function (paths relative to lemolatoon/ebi).
#pragma once

#include "codecs.hpp"

namespace ebi {

// ------------------------------------------------------------ bit helpers
__device__ __forceinline__ uint32_t mask_byte(const uint64_t *mask, uint64_t g, uint32_t cnt) {
  // bits of records [g, g + cnt) of the global bitset, packed LSB-first
  if (!mask) return (1u << cnt) - 1;
  const uint64_t w = g >> 6;
  const uint32_t sh = (uint32_t)(g & 63);
  uint64_t v = mask[w] >> sh;
  if (sh > 56) v |= mask[w + 1] << (64 - sh);
  return (uint32_t)v & ((1u << cnt) - 1);
}

// Write this lane's 8 result bits (records [g, g + 8) of one chunk; bits of
// records past the chunk end are 0).  A fully owned, byte-aligned group is
// stored as a byte; groups that share a byte with another chunk use 32-bit
// atomic ORs.  The output bitset is zeroed before the launch.
__device__ __forceinline__ void put_bits(uint64_t *out, uint64_t g, uint32_t bits, bool owned) {
  if (owned) {
    reinterpret_cast<uint8_t *>(out)[g >> 3] = (uint8_t)bits;
    return;
  }
  if (!bits) return;
  uint32_t *w = reinterpret_cast<uint32_t *>(out);
  const uint64_t wi = g >> 5;
  const uint32_t sh = (uint32_t)(g & 31);
  const uint64_t v = (uint64_t)bits << sh;
  if ((uint32_t)v) atomicOr(w + wi, (uint32_t)v);
  if ((uint32_t)(v >> 32)) atomicOr(w + wi + 1, (uint32_t)(v >> 32));
}

// The f64 predicate of a query, held by value in the hot loops.
struct PredF64 {
  int32_t kind, start_kind, end_kind;
  double value, start, end;
};
__device__ __forceinline__ PredF64 pred_of(const QueryArgs &a) {
  return PredF64{a.kind, a.start_kind, a.end_kind, a.value, a.start, a.end};
}

// Predicate::eval on f64 (decoder/query.rs:374-433).
template <class A>
__device__ __forceinline__ bool pred_f64(const A &a, double v) {
  if (a.kind == 0) return v == a.value;
  if (a.kind == 1) return v != a.value;
  bool s = true, e = true;
  if (a.start_kind == 1) s = v >= a.start;
  else if (a.start_kind == 2) s = v > a.start;
  if (a.end_kind == 1) e = v <= a.end;
  else if (a.end_kind == 2) e = v < a.end;
  return s && e;
}

// Sprintz predicate in the quantized domain (chunk_reader/sprintz.rs:116):
// lmode/rmode carry the bound kinds, ldp/rdp the quantized operands.
__device__ __forceinline__ bool pred_q(const QueryArgs &a, const ChunkPred &p, int64_t q) {
  if (a.kind == 0) return q == (int64_t)p.ldp;
  if (a.kind == 1) return q != (int64_t)p.ldp;
  bool s = true, e = true;
  if (p.lmode == 1) s = q >= (int64_t)p.ldp;
  else if (p.lmode == 2) s = q > (int64_t)p.ldp;
  if (p.rmode == 1) e = q <= (int64_t)p.rdp;
  else if (p.rmode == 2) e = q < (int64_t)p.rdp;
  return s && e;
}

// BUFF filter on one record (decoder/chunk_reader/buff.rs:128-228 with
// buff/filter_cmp.rs:298-451 and buff/filter_eq.rs:73-201).  `m` is the
// record's bitmask bit (1 if no bitmask), `has_mask` whether one was given.
__device__ __forceinline__ bool buff_pred(const QueryArgs &a, const ChunkPred &p, uint64_t d, bool m) {
  if (a.kind <= 1) {
    bool eq;
    if (p.lmode == kSideAll) eq = true;           // B == 0 and pred == base: all(), mask ignored
    else if (p.lmode == kSideEmpty) eq = false;
    else eq = m && (p.eq_u8 ? d == (p.ldp & 0xff) : d == p.ldp);
    return a.kind == 0 ? eq : (m && !eq);
  }
  bool l = true, r = true;
  const bool has_l = p.lmode != kSideNone, has_r = p.rmode != kSideNone;
  if (has_l) {
    if (p.lmode == kSideAll) l = true;
    else if (p.lmode == kSideEmpty) l = false;
    else l = m && (d > p.ldp || (p.lincl && d == p.ldp));
  }
  if (has_r) {
    if (p.rmode == kSideAll) r = true;
    else if (p.rmode == kSideEmpty) r = false;
    else r = m && (d < p.rdp || (p.rincl && d == p.rdp));
  }
  if (!has_l && !has_r) return m;
  return l && r;
}

// ---------------------------------------------------------------- filters
// mode 0: select (bit = mask); 1: predicate, in the code domain for BUFF and
// DeltaSprintz; 2: predicate on decoded f64 records.  One wave per record tile.
__global__ void __launch_bounds__(256) k_filter(const ColView *cols, const Job *jobs, uint32_t n_jobs,
                                               QueryArgs a, int mode) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_jobs) return;
  const Job jb = jobs[j];
  const ColView cv = cols[jb.col];
  if (!a.chunk_sel[jb.chunk]) return;
  const DevChunk c = cv.chunks[jb.chunk];
  const uint32_t lane = lane_id();
  const uint32_t i0 = jb.tile * kTile + lane * 8;
  // Sprintz needs every lane in the wave scan, so no early return before it.
  int64_t q[8];
  if (mode == 1 && c.codec == kCodecSprintz) sprintz_q(c, jb.tile, cv.tile_index[c.tile0 + jb.tile], q);
  if (i0 >= c.n) return;
  const uint32_t cnt = min(8u, c.n - i0);
  const uint64_t g = c.rec_off + i0;
  const uint32_t m = mask_byte(a.mask, g, cnt);
  uint32_t bits = 0;
  if (mode == 0) {
    bits = m;
  } else if (mode == 1 && c.codec == kCodecBUFF) {
    const ChunkPred p = a.cpred[jb.chunk];
    uint64_t d[8];
    buff_deltas(c, i0, d);
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((uint32_t)k < cnt && buff_pred(a, p, d[k], (m >> k) & 1)) bits |= 1u << k;
  } else if (mode == 1 && c.codec == kCodecSprintz) {
    const ChunkPred p = a.cpred[jb.chunk];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((uint32_t)k < cnt && ((m >> k) & 1) && pred_q(a, p, q[k])) bits |= 1u << k;
  } else {
    const double *v = a.values + g;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((uint32_t)k < cnt && ((m >> k) & 1) && pred_f64(a, v[k])) bits |= 1u << k;
  }
  put_bits(a.out_bits, g, bits, (g & 7) == 0 && cnt == 8);
}

// ------------------------------------------------------------- aggregates
// Order-preserving u64 keys.
__device__ __forceinline__ uint64_t f64_key(double x) {
  const uint64_t b = as_u64(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ uint64_t i64_key(int64_t x) { return (uint64_t)x ^ 0x8000000000000000ull; }

// Double-double addition (Knuth two-sum), deterministic for a fixed order.
__device__ __forceinline__ void dd_add(double &hi, double &lo, double bh, double bl) {
  const double s = hi + bh;
  if (!isfinite(s)) {  // infinities and NaN propagate as in a plain sum
    hi = s;
    lo = 0.0;
    return;
  }
  const double v = s - hi;
  const double e = (hi - (s - v)) + (bh - v);
  const double t = e + lo + bl;
  hi = s + t;
  lo = t - (hi - s);
}

// op: 0 sum, 1 min, 2 max.  Per-chunk accumulators acc[4*chunk + {sum, count,
// min key, max key}] hold exact integer sums for BUFF (plane deltas,
// buff/sum.rs:60-119) and DeltaSprintz (quantized q, sprintz.rs:267-351) and
// keys for min/max; f64 sums of the other codecs go to per-tile
// double-double partials.
__global__ void __launch_bounds__(256) k_aggregate(const ColView *cols, const Job *jobs, uint32_t n_jobs,
                                                  QueryArgs a) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_jobs) return;
  const Job jb = jobs[j];
  const ColView cv = cols[jb.col];
  if (!a.chunk_sel[jb.chunk]) return;
  const DevChunk c = cv.chunks[jb.chunk];
  const uint32_t lane = lane_id();
  const uint32_t i0 = jb.tile * kTile + lane * 8;
  const uint32_t cnt = i0 < c.n ? min(8u, c.n - i0) : 0u;
  const uint64_t g = c.rec_off + i0;
  const uint32_t m = cnt ? mask_byte(a.mask, g, cnt) : 0u;
  uint64_t *acc = a.acc + 4ull * jb.chunk;
  if (c.codec == kCodecBUFF || c.codec == kCodecSprintz) {
    uint64_t s = 0, n = 0, mn = ~0ull, mx = 0;
    if (c.codec == kCodecBUFF) {
      uint64_t d[8];
      buff_deltas(c, i0, d);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((m >> k) & 1) {
          s += d[k];
          ++n;
          mn = min(mn, d[k]);
          mx = max(mx, d[k]);
        }
    } else {
      int64_t q[8];
      sprintz_q(c, jb.tile, cv.tile_index[c.tile0 + jb.tile], q);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((m >> k) & 1) {
          s += (uint64_t)q[k];
          ++n;
          mn = min(mn, i64_key(q[k]));
          mx = max(mx, i64_key(q[k]));
        }
    }
    if (a.op == 0) {
      s = wave_sum_u64(s);
      n = wave_sum_u64(n);
      if (lane == 0 && n) {
        atomicAdd((unsigned long long *)acc, (unsigned long long)s);
        atomicAdd((unsigned long long *)acc + 1, (unsigned long long)n);
      }
    } else {
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        mn = min(mn, (uint64_t)__shfl_xor((unsigned long long)mn, d, kWave));
        mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, d, kWave));
      }
      if (lane == 0) {
        if (a.op == 1) atomicMin((unsigned long long *)acc + 2, (unsigned long long)mn);
        else atomicMax((unsigned long long *)acc + 3, (unsigned long long)mx);
      }
    }
    return;
  }
  // default codecs: decoded records
  const double *v = a.values + g;
  if (a.op == 0) {
    double hi = 0.0, lo = 0.0;
    for (int k = 0; k < 8; ++k)
      if ((m >> k) & 1) dd_add(hi, lo, v[k], 0.0);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const double oh = __shfl_xor(hi, d, kWave), ol = __shfl_xor(lo, d, kWave);
      // combine in a lane-independent order so every lane gets the same pair
      if ((lane & d) == 0) dd_add(hi, lo, oh, ol);
      else {
        double h2 = oh, l2 = ol;
        dd_add(h2, l2, hi, lo);
        hi = h2;
        lo = l2;
      }
    }
    if (lane == 0) {
      double *p = a.partials + 2ull * (c.rtile0 + jb.tile);
      p[0] = hi;
      p[1] = lo;
    }
  } else {
    uint64_t mn = ~0ull, mx = 0;
    for (int k = 0; k < 8; ++k)
      if (((m >> k) & 1) && !(v[k] != v[k])) {  // Rust f64::min/max ignore NaN
        mn = min(mn, f64_key(v[k]));
        mx = max(mx, f64_key(v[k]));
      }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      mn = min(mn, (uint64_t)__shfl_xor((unsigned long long)mn, d, kWave));
      mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, d, kWave));
    }
    if (lane == 0) {
      if (a.op == 1) atomicMin((unsigned long long *)acc + 2, (unsigned long long)mn);
      else atomicMax((unsigned long long *)acc + 3, (unsigned long long)mx);
    }
  }
}

// Per-chunk accumulator init: sum 0, count 0, min key max, max key 0.
__global__ void k_acc_init(uint64_t *acc, uint32_t n_chunks) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_chunks) return;
  acc[4ull * i] = 0;
  acc[4ull * i + 1] = 0;
  acc[4ull * i + 2] = ~0ull;
  acc[4ull * i + 3] = 0;
}

// One wave per chunk: lane l folds the chunk's tile partials l, l + 64,
// l + 128, ... (double-double), then a butterfly over the lanes; lane 0
// writes.  The order is fixed, so every call gives the same sum.
__global__ void __launch_bounds__(256) k_chunk_dd(const ColView *cols, const uint32_t *chunk_sel,
                                                 const double *partials, double *chunk_sum, uint32_t n_chunks) {
  const uint32_t ci = blockIdx.x * 4 + threadIdx.x / kWave;
  if (ci >= n_chunks) return;  // wave-uniform
  const uint32_t lane = lane_id();
  double hi = 0.0, lo = 0.0;
  if (chunk_sel[ci]) {  // wave-uniform
    const DevChunk c = cols[0].chunks[ci];
    const uint32_t nt = (c.n + kTile - 1) / kTile;
    for (uint32_t t = lane; t < nt; t += kWave)
      dd_add(hi, lo, partials[2ull * (c.rtile0 + t)], partials[2ull * (c.rtile0 + t) + 1]);
    for (int d = kWave / 2; d >= 1; d >>= 1) {
      const double oh = __shfl_xor(hi, d), ol = __shfl_xor(lo, d);
      dd_add(hi, lo, oh, ol);
    }
  }
  if (lane == 0) {
    chunk_sum[2ull * ci] = hi;
    chunk_sum[2ull * ci + 1] = lo;
  }
}

// Chunk selection: chunk_id filter and bitmask pruning (api/decoder.rs:216-226).
// One wave per chunk ORs the bitmask words of its record range.
__global__ void __launch_bounds__(256) k_chunk_select(const ColView *cols, const uint64_t *mask, int64_t chunk_id,
                                                     uint32_t *sel, uint32_t n_chunks) {
  const uint32_t ci = blockIdx.x * 4 + threadIdx.x / kWave;
  if (ci >= n_chunks) return;
  const uint32_t lane = lane_id();
  uint32_t s = (chunk_id < 0 || (uint64_t)chunk_id == ci) ? 1u : 0u;
  if (s && mask) {
    const DevChunk c = cols[0].chunks[ci];
    const uint64_t a = c.rec_off, b = c.rec_off + c.n;  // [a, b)
    uint64_t any = 0;
    for (uint64_t w = (a >> 6) + lane; w <= ((b - 1) >> 6) && b > a; w += kWave) {
      uint64_t x = mask[w];
      const uint64_t lo = w << 6;
      if (a > lo) x &= ~0ull << (a - lo);
      if (b < lo + 64) x &= (1ull << (b - lo)) - 1;
      any |= x;
    }
    s = __any(any != 0) ? 1u : 0u;
  }
  if (lane == 0) sel[ci] = s;
}

// --------------------------------------------------------------- compaction
// Selected records of `values` (bit r of `bits` set) are written in record
// order to `out`.  Block b covers words [b*4096, (b+1)*4096).
constexpr uint32_t kScanWords = 2;  // words per thread (a block: 512 words = 32,768 records)
constexpr uint32_t kScanBlock = 256;

__device__ __forceinline__ uint64_t block_excl_scan(uint64_t x, uint64_t *lds, uint64_t *total) {
  const uint32_t t = threadIdx.x, wv = t / kWave;
  uint64_t incl = x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t y = __shfl_up((unsigned long long)incl, d, kWave);
    if ((int)lane_id() >= d) incl += y;
  }
  if (lane_id() == kWave - 1) lds[wv] = incl;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (uint32_t k = 0; k < kScanBlock / kWave; ++k) {
    if (k < wv) off += lds[k];
    tot += lds[k];
  }
  __syncthreads();
  *total = tot;
  return off + incl - x;
}

__device__ __forceinline__ void popc_block(const uint64_t *bits, uint64_t n_words, uint64_t *block_cnt, uint32_t blk,
                                           uint64_t *lds) {
  const uint64_t w0 = ((uint64_t)blk * kScanBlock + threadIdx.x) * kScanWords;
  uint64_t c = 0;
  for (uint32_t k = 0; k < kScanWords; ++k)
    if (w0 + k < n_words) c += __popcll(bits[w0 + k]);
  uint64_t total;
  block_excl_scan(c, lds, &total);
  if (threadIdx.x == 0) block_cnt[blk] = total;
}

__global__ void __launch_bounds__(256) k_popc_blocks(const uint64_t *bits, uint64_t n_words, uint64_t *block_cnt) {
  __shared__ uint64_t lds[kScanBlock / kWave];
  popc_block(bits, n_words, block_cnt, blockIdx.x, lds);
}

// One block: exclusive scan of block counts in place; total to *n_out.
__device__ __forceinline__ void scan_counts(uint64_t *cnt, uint32_t n, uint64_t *n_out, uint64_t *lds) {
  uint64_t carry = 0;
  for (uint32_t base = 0; base < n; base += kScanBlock) {
    const uint32_t i = base + threadIdx.x;
    const uint64_t x = i < n ? cnt[i] : 0;
    uint64_t total;
    const uint64_t ex = block_excl_scan(x, lds, &total);
    if (i < n) cnt[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) *n_out = carry;
}

__global__ void __launch_bounds__(256) k_scan_counts(uint64_t *cnt, uint32_t n, uint64_t *n_out) {
  __shared__ uint64_t lds[kScanBlock / kWave];
  scan_counts(cnt, n, n_out, lds);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t src) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// Writes the selected records of the block's words to out[block_off + ...]
// in record order.  Each thread holds kScanWords bitset words; the wave then
// walks its threads' words one 64-record word at a time, lane l taking
// record l (its output slot = the word's first slot + the set bits below l),
// so the value loads and the output stores are contiguous across the wave;
// eight threads' words per round, loads issued before the stores.
__device__ __forceinline__ void scatter_block(const uint64_t *bits, uint64_t n_words, const uint64_t *block_off,
                                              const double *__restrict__ values, double *__restrict__ out, uint32_t blk,
                                              uint64_t *lds) {
  const uint64_t w0 = ((uint64_t)blk * kScanBlock + threadIdx.x) * kScanWords;
  uint64_t x[kScanWords];
  uint64_t c = 0;
#pragma unroll
  for (uint32_t k = 0; k < kScanWords; ++k) {
    x[k] = w0 + k < n_words ? bits[w0 + k] : 0;
    c += __popcll(x[k]);
  }
  uint64_t total;
  const uint64_t pos = block_off[blk] + block_excl_scan(c, lds, &total);
  const uint32_t lane = lane_id();
  const uint64_t wave_w0 = w0 - (uint64_t)lane * kScanWords;  // the wave's first word
  constexpr uint32_t kRound = 8;
  for (uint32_t s0 = 0; s0 < kWave; s0 += kRound) {
    double v[kRound][kScanWords];
    uint64_t dst[kRound][kScanWords];
    bool sel[kRound][kScanWords];
#pragma unroll
    for (uint32_t u = 0; u < kRound; ++u) {
      uint64_t p = readlane64(pos, s0 + u);
#pragma unroll
      for (uint32_t k = 0; k < kScanWords; ++k) {
        const uint64_t word = readlane64(x[k], s0 + u);
        const uint64_t w = wave_w0 + (uint64_t)(s0 + u) * kScanWords + k;
        sel[u][k] = (word >> lane) & 1;
        dst[u][k] = p + __builtin_amdgcn_mbcnt_hi((uint32_t)(word >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)word, 0));
        v[u][k] = sel[u][k] ? values[(w << 6) + lane] : 0.0;
        p += __popcll(word);
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < kRound; ++u)
#pragma unroll
      for (uint32_t k = 0; k < kScanWords; ++k)
        if (sel[u][k]) out[dst[u][k]] = v[u][k];
  }
}

__global__ void __launch_bounds__(256) k_scatter(const uint64_t *bits, uint64_t n_words, const uint64_t *block_off,
                                                const double *__restrict__ values, double *__restrict__ out) {
  __shared__ uint64_t lds[kScanBlock / kWave];
  scatter_block(bits, n_words, block_off, values, out, blockIdx.x, lds);
}

// ------------------------------------------------------------ group kernels
// The kernels of a group of query plans (host/plan.cpp): each launch covers
// every plan of the group, so a group costs a fixed number of launches
// whatever its size.  Flat work items map to (plan, item) through a prefix
// over the plans (prefix[p] = first item of plan p, prefix[n] = total).
__device__ __forceinline__ uint32_t plan_of(const uint32_t *prefix, uint32_t n_plans, uint32_t x) {
  uint32_t lo = 0, hi = n_plans;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= x) lo = mid; else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// Scratch fills (zeroed bitsets, accumulator patterns): block row y = fill y.
__global__ void __launch_bounds__(256) k_group_fill(const GroupFill *fills) {
  const GroupFill f = fills[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = f.period == 4 ? (uint32_t)(i & 3) : 0u;
    f.p[i] = r == 0 ? f.v[0] : r == 1 ? f.v[1] : r == 2 ? f.v[2] : f.v[3];
  }
}

// Chunk selection of every plan (chunk_id, bitmask pruning; api/decoder.rs:
// 216-226): one wave per (plan, chunk).
__global__ void __launch_bounds__(256) k_group_select(const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans,
                                                     uint32_t n_items) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_items) return;
  const uint32_t p = plan_of(prefix, n_plans, j);
  const PlanDesc &P = plans[p];
  const uint32_t ci = j - prefix[p];
  const uint64_t *mask = P.a.mask;
  uint32_t s = (P.chunk_id < 0 || (uint64_t)P.chunk_id == ci) ? 1u : 0u;
  if (s && mask) {
    const DevChunk c = P.view->chunks[ci];
    const uint64_t a = c.rec_off, b = c.rec_off + c.n;
    uint64_t any = 0;
    for (uint64_t w = (a >> 6) + lane_id(); b > a && w <= ((b - 1) >> 6); w += kWave) {
      uint64_t x = mask[w];
      const uint64_t lo = w << 6;
      if (a > lo) x &= ~0ull << (a - lo);
      if (b < lo + 64) x &= (1ull << (b - lo)) - 1;
      any |= x;
    }
    s = __any(any != 0) ? 1u : 0u;
  }
  if (lane_id() == 0) const_cast<uint32_t *>(P.a.chunk_sel)[ci] = s;
}

// One record tile of one plan: the filter (bits of the predicate, or of the
// bitmask when the plan only selects) and the aggregate, in one pass over
// the tile's codes or records -- BUFF planes and DeltaSprintz codes are read
// once even when the plan filters and then sums (FILTER_SUM).  The filter
// follows k_filter, the aggregate k_aggregate; with kPfAggSel the aggregate
// takes the predicate's selection instead of the bitmask (as sum(Some(
// filter(..))) does) and counts it per chunk (acc[1]) for the re-pruning.
__device__ __forceinline__ void group_tile(const PlanDesc &P, const Job jb) {
  const QueryArgs &a = P.a;
  if (!a.chunk_sel[jb.chunk]) return;  // wave-uniform
  const ColView *cv = P.view;
  const DevChunk c = cv->chunks[jb.chunk];
  const uint32_t lane = lane_id();
  const uint32_t i0 = jb.tile * kTile + lane * 8;
  const uint32_t cnt = i0 < c.n ? min(8u, c.n - i0) : 0u;
  const uint64_t g = c.rec_off + i0;
  const uint32_t m = cnt ? mask_byte(a.mask, g, cnt) : 0u;
  const uint32_t flags = P.flags;
  const int agg = (int)((flags >> kPfAggShift) & 3u);
  const bool pred = (flags & kPfPred) != 0;
  const bool insitu = c.codec == kCodecBUFF || c.codec == kCodecSprintz;
  const bool code = insitu && ((pred && P.filter_mode == 1) || agg);
  uint64_t d[8];
  int64_t q[8];
  if (code && c.codec == kCodecSprintz) sprintz_q(c, jb.tile, cv->tile_index[c.tile0 + jb.tile], q);  // every lane
  if (code && c.codec == kCodecBUFF) buff_deltas(c, i0, d);  // zeros past the chunk's end
  uint32_t sel = m;
  if (pred) {
    uint32_t bits = 0;
    if (P.filter_mode == 1 && c.codec == kCodecBUFF) {
      const ChunkPred cp = a.cpred[jb.chunk];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((uint32_t)k < cnt && buff_pred(a, cp, d[k], (m >> k) & 1)) bits |= 1u << k;
    } else if (P.filter_mode == 1 && c.codec == kCodecSprintz) {
      const ChunkPred cp = a.cpred[jb.chunk];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((uint32_t)k < cnt && ((m >> k) & 1) && pred_q(a, cp, q[k])) bits |= 1u << k;
    } else if (P.filter_mode == 0) {
      bits = m;
    } else {
      const double *v = a.values + g;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((uint32_t)k < cnt && ((m >> k) & 1) && pred_f64(a, v[k])) bits |= 1u << k;
    }
    sel = bits;
  }
  if ((flags & kPfBits) && cnt) put_bits(a.out_bits, g, sel, (g & 7) == 0 && cnt == 8);
  if (!agg) return;  // wave-uniform
  const uint32_t use = (flags & kPfAggSel) ? sel : m;
  uint64_t *acc = a.acc + 4ull * jb.chunk;
  if (insitu) {
    uint64_t sum = 0, n = 0, mn = ~0ull, mx = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((use >> k) & 1) {
        const uint64_t x = c.codec == kCodecBUFF ? d[k] : (uint64_t)q[k];
        const uint64_t key = c.codec == kCodecBUFF ? d[k] : i64_key(q[k]);
        sum += x;
        ++n;
        mn = min(mn, key);
        mx = max(mx, key);
      }
    if (agg == 1) {
      sum = wave_sum_u64(sum);
      n = wave_sum_u64(n);
      if (lane == 0 && n) {
        atomicAdd((unsigned long long *)acc, (unsigned long long)sum);
        atomicAdd((unsigned long long *)acc + 1, (unsigned long long)n);
      }
    } else {
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        mn = min(mn, (uint64_t)__shfl_xor((unsigned long long)mn, o, kWave));
        mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, o, kWave));
      }
      if (lane == 0) {
        if (agg == 2) atomicMin((unsigned long long *)acc + 2, (unsigned long long)mn);
        else atomicMax((unsigned long long *)acc + 3, (unsigned long long)mx);
      }
    }
    return;
  }
  const double *v = a.values + g;
  if (agg == 1) {
    double hi = 0.0, lo = 0.0;
    uint32_t n = 0;
    for (int k = 0; k < 8; ++k)
      if ((use >> k) & 1) {
        dd_add(hi, lo, v[k], 0.0);
        ++n;
      }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double oh = __shfl_xor(hi, o, kWave), ol = __shfl_xor(lo, o, kWave);
      if ((lane & o) == 0) {
        dd_add(hi, lo, oh, ol);
      } else {
        double h2 = oh, l2 = ol;
        dd_add(h2, l2, hi, lo);
        hi = h2;
        lo = l2;
      }
    }
    if (flags & kPfAggSel) n = (uint32_t)wave_sum_u64(n);
    if (lane == 0) {
      double *pp = a.partials + 2ull * (c.rtile0 + jb.tile);
      pp[0] = hi;
      pp[1] = lo;
      if ((flags & kPfAggSel) && n) atomicAdd((unsigned long long *)acc + 1, (unsigned long long)n);
    }
  } else {
    uint64_t mn = ~0ull, mx = 0;
    for (int k = 0; k < 8; ++k)
      if (((use >> k) & 1) && !(v[k] != v[k])) {  // Rust f64::min/max ignore NaN
        mn = min(mn, f64_key(v[k]));
        mx = max(mx, f64_key(v[k]));
      }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      mn = min(mn, (uint64_t)__shfl_xor((unsigned long long)mn, o, kWave));
      mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, o, kWave));
    }
    if (lane == 0) {
      if (agg == 2) atomicMin((unsigned long long *)acc + 2, (unsigned long long)mn);
      else atomicMax((unsigned long long *)acc + 3, (unsigned long long)mx);
    }
  }
}

// The record-tile pass of every plan with kPfTiles: one wave per (plan,
// record tile), plans' tile jobs concatenated.
__global__ void __launch_bounds__(256) k_group_tiles(const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans,
                                                    uint32_t n_jobs) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_jobs) return;
  const uint32_t p = plan_of(prefix, n_plans, j);
  const PlanDesc &P = plans[p];
  group_tile(P, P.rec_jobs[j - prefix[p]]);
}

// Rust f64::min / f64::max: a NaN operand yields the other operand.
__device__ __forceinline__ double rmin_d(double a, double b) { return a != a ? b : b != b ? a : (b < a ? b : a); }
__device__ __forceinline__ double rmax_d(double a, double b) { return a != a ? b : b != b ? a : (b > a ? b : a); }

// Per-chunk results of every plan with kPfFold, one wave per (plan, chunk):
// whether the Decoder's loop takes the chunk (selected; for FILTER_SUM also
// holding a selected record -- sum(Some(filtered)) prunes the others) into
// chunk_sel, and what it adds: an f64 sum's double-double over the chunk's
// tile partials (a fixed order), else the chunk's value by the codec
// reader's rules (BUFF buff/sum.rs:60-119, min_max.rs:42-341 with the B = 0
// and unmasked-min exits; DeltaSprintz sprintz.rs:176-446, NaN for an empty
// in-chunk bitmask; default readers' f64::min / max keys).
__global__ void __launch_bounds__(256) k_group_fold(const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans,
                                                   uint32_t n_items) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_items) return;
  const uint32_t p = plan_of(prefix, n_plans, j);
  const PlanDesc &P = plans[p];
  if (!(P.flags & kPfFold)) return;  // wave-uniform
  const uint32_t ci = j - prefix[p];
  const uint64_t *ak = P.a.acc + 4ull * ci;
  uint32_t take = P.a.chunk_sel[ci];
  if (take && (P.flags & kPfAggSel)) take = ak[1] != 0;
  const int agg = (int)((P.flags >> kPfAggShift) & 3u);  // 1 sum, 2 min, 3 max
  double *cs = P.chunk_sum + 2ull * ci;
  if (P.flags & kPfF64Sum) {
    double hi = 0.0, lo = 0.0;
    if (take) {  // wave-uniform
      const DevChunk c = P.view->chunks[ci];
      const uint32_t nt = (c.n + kTile - 1) / kTile;
      for (uint32_t t = lane_id(); t < nt; t += kWave)
        dd_add(hi, lo, P.a.partials[2ull * (c.rtile0 + t)], P.a.partials[2ull * (c.rtile0 + t) + 1]);
      for (int d = kWave / 2; d >= 1; d >>= 1) {
        const double oh = __shfl_xor(hi, d), ol = __shfl_xor(lo, d);
        dd_add(hi, lo, oh, ol);
      }
    }
    if (lane_id() == 0) {
      cs[0] = hi;
      cs[1] = lo;
      const_cast<uint32_t *>(P.a.chunk_sel)[ci] = take;
    }
    return;
  }
  if (lane_id() != 0) return;
  double v = 0.0;
  if (take) {
    const DevChunk c = P.view->chunks[ci];
    const bool mask = (P.flags & (kPfMask | kPfAggSel)) != 0;
    if (c.codec == kCodecBUFF) {
      const double scale = ldexp(1.0, (int)c.frac);  // 2.0f64.powi(frac)
      if (agg == 1) {
        if (c.bits == 0)  // buff/sum.rs:80-84: value * n, bitmask ignored
          v = ((double)c.base / scale) * (double)c.n;
        else
          v = (double)(int64_t)(ak[0] + (uint64_t)c.base * ak[1]) / scale;
      } else if (c.bits == 0 || (agg == 2 && !mask)) {
        v = (double)c.base / scale;  // min_max.rs:316-341 and the B = 0 exits
      } else {
        v = (double)(int64_t)((uint64_t)c.base + (agg == 2 ? ak[2] : ak[3])) / scale;
      }
    } else if (c.codec == kCodecSprintz) {
      const double sc = (double)c.scale;
      if (mask && ak[1] == 0 && agg == 1) v = as_f64(kNan);
      else if (agg == 1) v = (double)(int64_t)ak[0] / sc;
      else v = (double)(int64_t)((agg == 2 ? ak[2] : ak[3]) ^ 0x8000000000000000ull) / sc;
    } else {
      const uint64_t key = agg == 2 ? ak[2] : ak[3];
      if ((agg == 2 && key == ~0ull) || (agg == 3 && key == 0)) {
        v = agg == 2 ? INFINITY : -INFINITY;
      } else {
        v = as_f64((key >> 63) ? (key & 0x7fffffffffffffffull) : ~key);
      }
    }
  }
  cs[0] = v;
  cs[1] = 0.0;
  const_cast<uint32_t *>(P.a.chunk_sel)[ci] = take;
}

// The Decoder's fold (api/decoder.rs:323-439) of every plan with kPfFold,
// one block per plan: the chunks' results, staged through LDS, folded in
// chunk order by one lane -- `result += v` (the reference's f64 order), the
// double-double total of an f64 sum, f64::min / f64::max from +-INFINITY.
constexpr uint32_t kFoldStage = 2048;
__global__ void __launch_bounds__(256) k_group_finish(const PlanDesc *plans) {
  __shared__ double vs[2 * kFoldStage];
  __shared__ uint32_t take[kFoldStage];
  const PlanDesc &P = plans[blockIdx.x];
  if (!(P.flags & kPfFold)) return;  // block-uniform
  const int agg = (int)((P.flags >> kPfAggShift) & 3u);
  const bool f64sum = (P.flags & kPfF64Sum) != 0;
  double result = agg == 1 ? 0.0 : agg == 2 ? INFINITY : -INFINITY;
  double hi = 0.0, lo = 0.0;
  for (uint32_t k0 = 0; k0 < P.n_chunks; k0 += kFoldStage) {
    const uint32_t m = min(kFoldStage, P.n_chunks - k0);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
      take[i] = P.a.chunk_sel[k0 + i];
      vs[2 * i] = P.chunk_sum[2ull * (k0 + i)];
      vs[2 * i + 1] = P.chunk_sum[2ull * (k0 + i) + 1];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (uint32_t i = 0; i < m; ++i) {
        if (!take[i]) continue;
        if (f64sum) {  // keep full precision across chunks (two-sum)
          for (int t = 0; t < 2; ++t) {
            const double b = vs[2 * i + t];
            const double sum = hi + b;
            if (!isfinite(sum)) {
              hi = sum;
              lo = 0.0;
              continue;
            }
            const double bv = sum - hi;
            const double e = (hi - (sum - bv)) + (b - bv);
            hi = sum;
            lo += e;
          }
        } else if (agg == 1) {
          result += vs[2 * i];
        } else if (agg == 2) {
          result = rmin_d(result, vs[2 * i]);
        } else {
          result = rmax_d(result, vs[2 * i]);
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) P.result->scalar = f64sum ? (isfinite(hi) ? hi + lo : hi) : result;
}

// Stream compaction of every plan with kPfCompact (materialize through a
// bitset): per-block counts, a per-plan scan (its total to result->n_out),
// then the scatter; blocks of all plans concatenated (prefix over plans).
__global__ void __launch_bounds__(256) k_group_popc(const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans,
                                                   uint32_t n_blocks) {
  __shared__ uint64_t lds[kScanBlock / kWave];
  const uint32_t p = plan_of(prefix, n_plans, blockIdx.x);
  const PlanDesc &P = plans[p];
  if (!(P.flags & kPfCompact)) return;  // block-uniform
  popc_block(P.a.out_bits, P.words, P.block, blockIdx.x - prefix[p], lds);
}
__global__ void __launch_bounds__(256) k_group_scan(const PlanDesc *plans) {
  __shared__ uint64_t lds[kScanBlock / kWave];
  const PlanDesc &P = plans[blockIdx.x];
  if (!(P.flags & kPfCompact)) return;
  scan_counts(P.block, P.n_blocks, &P.result->n_out, lds);
}
__global__ void __launch_bounds__(256) k_group_scatter(const PlanDesc *plans, const uint32_t *prefix,
                                                      uint32_t n_plans, uint32_t n_blocks) {
  __shared__ uint64_t lds[kScanBlock / kWave];
  const uint32_t p = plan_of(prefix, n_plans, blockIdx.x);
  const PlanDesc &P = plans[p];
  if (!(P.flags & kPfCompact)) return;
  scatter_block(P.a.out_bits, P.words, P.block, P.a.values, P.out, blockIdx.x - prefix[p], lds);
}

namespace launch {

void group_fill(hipStream_t s, const GroupFill *fills, uint32_t n_fills, uint64_t max_words) {
  if (!n_fills) return;
  const uint32_t bx = (uint32_t)std::min<uint64_t>((max_words + 255) / 256, 1024);
  hipLaunchKernelGGL(k_group_fill, dim3(std::max(bx, 1u), n_fills), dim3(256), 0, s, fills);
}
void group_select(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_items) {
  if (n_items)
    hipLaunchKernelGGL(k_group_select, dim3((n_items + 3) / 4), dim3(256), 0, s, plans, prefix, n_plans, n_items);
}
void group_tiles(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_jobs) {
  if (n_jobs) hipLaunchKernelGGL(k_group_tiles, dim3((n_jobs + 3) / 4), dim3(256), 0, s, plans, prefix, n_plans, n_jobs);
}
void group_fold(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_items,
                bool any) {
  if (!any) return;
  if (n_items)
    hipLaunchKernelGGL(k_group_fold, dim3((n_items + 3) / 4), dim3(256), 0, s, plans, prefix, n_plans, n_items);
  hipLaunchKernelGGL(k_group_finish, dim3(n_plans), dim3(256), 0, s, plans);
}
void group_compact(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans,
                   uint32_t n_blocks) {
  if (!n_blocks) return;
  hipLaunchKernelGGL(k_group_popc, dim3(n_blocks), dim3(kScanBlock), 0, s, plans, prefix, n_plans, n_blocks);
  hipLaunchKernelGGL(k_group_scan, dim3(n_plans), dim3(kScanBlock), 0, s, plans);
  hipLaunchKernelGGL(k_group_scatter, dim3(n_blocks), dim3(kScanBlock), 0, s, plans, prefix, n_plans, n_blocks);
}

void chunk_select(hipStream_t s, const ColView *cols, const uint64_t *mask, int64_t chunk_id, uint32_t *sel,
                  uint32_t n_chunks) {
  if (n_chunks)
    hipLaunchKernelGGL(k_chunk_select, dim3((n_chunks + 3) / 4), dim3(256), 0, s, cols, mask, chunk_id, sel, n_chunks);
}

void filter(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, const QueryArgs &a, int mode) {
  if (n_jobs) hipLaunchKernelGGL(k_filter, dim3((n_jobs + 3) / 4), dim3(256), 0, s, cols, jobs, n_jobs, a, mode);
}

void aggregate(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, const QueryArgs &a,
               uint32_t n_chunks) {
  if (n_chunks) hipLaunchKernelGGL(k_acc_init, dim3((n_chunks + 255) / 256), dim3(256), 0, s, a.acc, n_chunks);
  if (n_jobs) hipLaunchKernelGGL(k_aggregate, dim3((n_jobs + 3) / 4), dim3(256), 0, s, cols, jobs, n_jobs, a);
}

void chunk_dd(hipStream_t s, const ColView *cols, const uint32_t *sel, const double *partials, double *chunk_sum,
              uint32_t n_chunks) {
  if (n_chunks)
    hipLaunchKernelGGL(k_chunk_dd, dim3((n_chunks + 3) / 4), dim3(256), 0, s, cols, sel, partials, chunk_sum,
                       n_chunks);
}

uint32_t compact_blocks(uint64_t n_words) {
  const uint64_t per = (uint64_t)kScanBlock * kScanWords;
  return (uint32_t)((n_words + per - 1) / per);
}

void compact(hipStream_t s, const uint64_t *bits, uint64_t n_words, const double *values, double *out,
             uint64_t *block_scratch, uint64_t *n_out) {
  const uint32_t nb = compact_blocks(n_words);
  if (!nb) return;
  hipLaunchKernelGGL(k_popc_blocks, dim3(nb), dim3(kScanBlock), 0, s, bits, n_words, block_scratch);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(kScanBlock), 0, s, block_scratch, nb, n_out);
  hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(kScanBlock), 0, s, bits, n_words, block_scratch, values, out);
}

}  // namespace launch
}  // namespace ebi
