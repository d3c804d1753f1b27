// Query kernels of libebi_hip (gfx950): filters to a bitset, aggregates and
// stream compaction.  Included once, by decode.hip (one device TU).
//
// Work unit: a record tile (kTile = 512 records of one chunk, 8 consecutive
// records per lane).  BUFF and DeltaSprintz evaluate predicates and
// aggregates in their code domain (plane deltas, quantized q); the other
// codecs work on records decoded into HBM.  Semantics follow the reference
// readers cited on each function (paths relative to lemolatoon/ebi).
#pragma once

#include "../chunk_pred.hpp"
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

// The group's first launch, two independent parts in one 1-D grid:
//  * blocks [0, n_fills * fbx): scratch fills (zeroed bitsets, accumulator
//    patterns), fbx blocks per fill;
//  * the rest: chunk selection of every plan (chunk_id, bitmask pruning;
//    api/decoder.rs:216-226), one wave per (plan, chunk), which also writes
//    the chunk's in-situ predicate parameters (ChunkPred) for kPfCpred plans.
__global__ void __launch_bounds__(256) k_group_prep(const GroupFill *fills, uint32_t n_fills, uint32_t fbx,
                                                   const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans,
                                                   uint32_t n_items) {
  if (blockIdx.x < n_fills * fbx) {  // block-uniform
    const GroupFill f = fills[blockIdx.x / fbx];
    const uint64_t step = (uint64_t)fbx * blockDim.x;
    for (uint64_t i = (uint64_t)(blockIdx.x % fbx) * blockDim.x + threadIdx.x; i < f.n; i += step) {
      const uint32_t r = f.period == 4 ? (uint32_t)(i & 3) : 0u;
      f.p[i] = r == 0 ? f.v[0] : r == 1 ? f.v[1] : r == 2 ? f.v[2] : f.v[3];
    }
    return;
  }
  const uint32_t j = (blockIdx.x - n_fills * fbx) * 4 + threadIdx.x / kWave;
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
  if (lane_id() != 0) return;
  const_cast<uint32_t *>(P.a.chunk_sel)[ci] = s;
  if (P.flags & kPfCpred) {
    const DevChunk c = P.view->chunks[ci];
    const ebi_pred pr{P.a.kind, P.a.start_kind, P.a.end_kind, 0, P.a.value, P.a.start, P.a.end};
    const_cast<ChunkPred *>(P.a.cpred)[ci] =
        P.scheme == kCodecBUFF ? buff_chunk_pred(c, pr, mask != nullptr) : sprintz_chunk_pred(c, pr);
  }
}

// ------------------------------------------------------- record-tile spans
// The record-tile pass works on spans: up to kSpanTiles consecutive tiles of
// one chunk per wave (the column's span jobs, Job::kind = tile count).  The
// wave loads the chunk's descriptor and predicate once, issues the loads of
// tile k + 1 before it works on tile k (two tiles in flight per wave), keeps
// its aggregates in lane registers across the span and adds them to the
// chunk's accumulators once per span.
//
// Decoding BUFF and DeltaSprintz codes costs few bytes per record, so these
// kernels are bound by vector ALU work rather than HBM: the predicate is
// folded once per chunk into an interval test on the code (CodeSel), only
// the requested aggregate is kept, and BUFF chunks of B < 24 bits work on
// 32-bit deltas.

// Lane-local aggregates of one wave over a span.
struct SpanAcc {
  uint64_t sum = 0, n = 0, mn = ~0ull, mx = 0;
};

// Which records of a chunk a code plan selects, from the code x of each
// record (BUFF delta, unsigned; DeltaSprintz quantized value, signed):
// none: no record; otherwise the record's bitmask bit (use_mask) and,
// unless all, ((x - lo) <= (hi - lo) as u64) != inv -- an interval test that
// holds for signed codes too (lo <= hi).
struct CodeSel {
  uint64_t lo, hi;
  bool none, use_mask, all, inv;
};

// buff_pred (chunk_reader/buff.rs:128-228, filter_cmp.rs:298-451,
// filter_eq.rs:73-201) folded per chunk: the early exits (kSideAll: every
// record, bitmask ignored; kSideEmpty: none) and the comparisons as bounds.
__device__ __forceinline__ CodeSel code_sel_buff(const PlanDesc &P, const ChunkPred &p) {
  CodeSel s{0, ~0ull, false, true, true, false};
  if (!(P.flags & kPfPred) || P.filter_mode != 1) return s;  // the bitmask alone
  if (P.a.kind <= 1) {
    const uint64_t v = p.eq_u8 ? (p.ldp & 0xff) : p.ldp;
    if (P.a.kind == 0) {  // Eq: all(), none, or m && d == v
      if (p.lmode == kSideAll) s.use_mask = false;
      else if (p.lmode == kSideEmpty) s.none = true;
      else s = CodeSel{v, v, false, true, false, false};
    } else {  // NotEq = m && !eq: none, m, or m && d != v
      if (p.lmode == kSideAll) s.none = true;
      else if (p.lmode != kSideEmpty) s = CodeSel{v, v, false, true, false, true};
    }
    return s;
  }
  const bool cl = p.lmode == kSideCmp, cr = p.rmode == kSideCmp;
  if (p.lmode == kSideNone && p.rmode == kSideNone) return s;  // m
  if (p.lmode == kSideEmpty || p.rmode == kSideEmpty) {
    s.none = true;
    return s;
  }
  s.use_mask = cl || cr;  // kSideAll on both sides, or one and None: every record
  s.all = !(cl || cr);
  if (cl) {
    if (!p.lincl && p.ldp == ~0ull) s.none = true;
    s.lo = p.lincl ? p.ldp : p.ldp + 1;
  }
  if (cr) {
    if (!p.rincl && p.rdp == 0) s.none = true;
    s.hi = p.rincl ? p.rdp : p.rdp - 1;
  }
  if (s.lo > s.hi) s.none = true;
  return s;
}

// pred_q (chunk_reader/sprintz.rs:116) as bounds on the signed quantized
// value; the bitmask always applies.
__device__ __forceinline__ CodeSel code_sel_sprintz(const PlanDesc &P, const ChunkPred &p) {
  CodeSel s{0x8000000000000000ull, 0x7fffffffffffffffull, false, true, true, false};
  if (!(P.flags & kPfPred) || P.filter_mode != 1) return s;
  if (P.a.kind <= 1) {
    s = CodeSel{p.ldp, p.ldp, false, true, false, P.a.kind == 1};
    return s;
  }
  if (p.lmode == 0 && p.rmode == 0) return s;
  int64_t lo = INT64_MIN, hi = INT64_MAX;
  if (p.lmode == 1) lo = (int64_t)p.ldp;
  else if (p.lmode == 2) {
    if ((int64_t)p.ldp == INT64_MAX) s.none = true;
    lo = (int64_t)(p.ldp + 1);
  }
  if (p.rmode == 1) hi = (int64_t)p.rdp;
  else if (p.rmode == 2) {
    if ((int64_t)p.rdp == INT64_MIN) s.none = true;
    hi = (int64_t)(p.rdp - 1);
  }
  if (lo > hi) s.none = true;
  s.lo = (uint64_t)lo;
  s.hi = (uint64_t)hi;
  s.all = false;
  return s;
}

// The records of this lane in tile t: count, global id of the first, and
// the bitmask bits (1 if no bitmask).
struct TileLane {
  uint32_t cnt, valid, m;
  uint64_t g;
};
__device__ __forceinline__ TileLane tile_lane(const PlanDesc &P, const DevChunk &c, uint32_t t) {
  TileLane L;
  const uint32_t i0 = t * kTile + lane_id() * 8;
  L.cnt = i0 < c.n ? min(8u, c.n - i0) : 0u;
  L.valid = 0xffu >> (8 - L.cnt);
  L.g = c.rec_off + i0;
  L.m = L.cnt ? mask_byte(P.a.mask, L.g, L.cnt) : 0u;
  return L;
}

// One record tile of a code plan: the filter (predicate or bitmask bits),
// then the one aggregate the plan asks for (the bitmask's records, or with
// kPfAggSel the predicate's -- as sum(Some(filter(..))) does -- counted for
// the re-pruning).  X: uint32_t (BUFF, B < 24) or uint64_t codes.
template <int KIND, class X>
__device__ __forceinline__ void code_tile(const PlanDesc &P, const CodeSel &cs, const TileLane &L, const X x[8],
                                          SpanAcc &acc) {
  uint32_t sel = cs.use_mask ? L.m : L.valid;
  if (!cs.all) {
    // 32-bit codes (B < 24) lie below 2^24: bounds above 2^32 - 1 clamp to it
    constexpr uint64_t top = sizeof(X) == 4 ? 0xffffffffull : ~0ull;
    const X lo = (X)(cs.lo < top ? cs.lo : top);
    const X span = (X)(cs.hi < top ? cs.hi : top) - lo;
    uint32_t in = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) in |= (X)(x[k] - lo) <= span ? 1u << k : 0u;
    sel &= cs.inv ? ~in : in;
  }
  const uint32_t flags = P.flags;
  if ((flags & kPfBits) && L.cnt) put_bits(P.a.out_bits, L.g, sel, (L.g & 7) == 0 && L.cnt == 8);
  const int agg = (int)((flags >> kPfAggShift) & 3u);
  if (agg == 1) {
    X s = 0;  // B < 24: 8 deltas sum below 2^27
#pragma unroll
    for (int k = 0; k < 8; ++k) s += ((sel >> k) & 1) ? x[k] : (X)0;
    acc.sum += (uint64_t)s;
    acc.n += (uint64_t)__builtin_popcount(sel);
  } else if (agg && KIND == kSpanBuff) {  // keys: the deltas themselves
    X mn = (X)~0ull, mx = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const bool on = (sel >> k) & 1;
      if (agg == 2) mn = min(mn, on ? x[k] : (X)~0ull);
      else mx = max(mx, on ? x[k] : (X)0);
    }
    if (agg == 2 && sel) acc.mn = min(acc.mn, (uint64_t)mn);
    if (agg == 3) acc.mx = max(acc.mx, (uint64_t)mx);
  } else if (agg) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((sel >> k) & 1) {
        const uint64_t key = i64_key((int64_t)x[k]);
        if (agg == 2) acc.mn = min(acc.mn, key);
        else acc.mx = max(acc.mx, key);
      }
  }
}

// One record tile of a DeltaSprintz plan with W <= 16: record j of the lane
// has q = base + s[j] (mod 2^64), s[j] an exact 32-bit prefix below 2^21 in
// magnitude, so the interval test, the sum and min / max work on s and the
// lane's base.  Unless base lies within 2^26 of the i64 range's ends (then
// per record, modulo 2^64), q = base + s exactly: lo <= q <= hi is
// lo - base <= s <= hi - base with both bounds clamped to 32 bits; the
// selected q sum to count * base + sum(s) (mod 2^64); min / max of base + s
// are base + min / max of s.
__device__ __forceinline__ int32_t clamp_i32(int64_t x) {
  return x < INT32_MIN ? INT32_MIN : x > INT32_MAX ? INT32_MAX : (int32_t)x;
}
__device__ __forceinline__ void sprintz_tile(const PlanDesc &P, const CodeSel &cs, const TileLane &L, int64_t base,
                                             const int32_t s[8], SpanAcc &acc) {
  constexpr int64_t kEdge = 1ll << 26;
  const bool inner = base > INT64_MIN + kEdge && base < INT64_MAX - kEdge;
  uint32_t sel = cs.use_mask ? L.m : L.valid;
  if (!cs.all) {
    uint32_t in = 0;
    if (inner) {
      // lo - base and hi - base, saturated: |base| < 2^63 - 2^26 keeps one side in range
      int64_t a, b;
      if (__builtin_sub_overflow((int64_t)cs.lo, base, &a)) a = base > 0 ? INT64_MIN : INT64_MAX;
      if (__builtin_sub_overflow((int64_t)cs.hi, base, &b)) b = base > 0 ? INT64_MIN : INT64_MAX;
      const int32_t a32 = clamp_i32(a), b32 = clamp_i32(b);
#pragma unroll
      for (int k = 0; k < 8; ++k) in |= (s[k] >= a32 && s[k] <= b32) ? 1u << k : 0u;  // |s| < 2^21
    } else {
      const uint64_t lo = cs.lo - (uint64_t)base, span = cs.hi - cs.lo;
#pragma unroll
      for (int k = 0; k < 8; ++k) in |= (uint64_t)(int64_t)s[k] - lo <= span ? 1u << k : 0u;
    }
    sel &= cs.inv ? ~in : in;
  }
  const uint32_t flags = P.flags;
  if ((flags & kPfBits) && L.cnt) put_bits(P.a.out_bits, L.g, sel, (L.g & 7) == 0 && L.cnt == 8);
  const int agg = (int)((flags >> kPfAggShift) & 3u);
  if (agg == 1) {
    int32_t t = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) t += ((sel >> k) & 1) ? s[k] : 0;
    const uint32_t n = (uint32_t)__builtin_popcount(sel);
    acc.sum += (uint64_t)n * (uint64_t)base + (uint64_t)(int64_t)t;
    acc.n += n;
  } else if (agg && sel) {
    if (inner) {
      int32_t m = agg == 2 ? INT32_MAX : INT32_MIN;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const bool on = (sel >> k) & 1;
        if (agg == 2) m = min(m, on ? s[k] : INT32_MAX);
        else m = max(m, on ? s[k] : INT32_MIN);
      }
      const uint64_t key = i64_key(base + m);
      if (agg == 2) acc.mn = min(acc.mn, key);
      else acc.mx = max(acc.mx, key);
    } else {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((sel >> k) & 1) {
          const uint64_t key = i64_key((int64_t)((uint64_t)base + (uint64_t)(int64_t)s[k]));
          if (agg == 2) acc.mn = min(acc.mn, key);
          else acc.mx = max(acc.mx, key);
        }
    }
  }
}

// One record tile of a plan on decoded records (v) or on the bitmask alone
// (filter_mode 0, no aggregate): the filter follows k_filter, the aggregate
// k_aggregate (f64 sums as a double-double partial per tile, folded in tile
// order later; min / max as order-preserving keys, NaN ignored as Rust's
// f64::min / max do).
__device__ __forceinline__ void values_tile(const PlanDesc &P, const DevChunk &c, uint32_t t, const double v[8],
                                            SpanAcc &acc) {
  const QueryArgs &a = P.a;
  const TileLane L = tile_lane(P, c, t);
  const uint32_t flags = P.flags;
  const int agg = (int)((flags >> kPfAggShift) & 3u);
  uint32_t sel = L.m;
  if (flags & kPfPred) {
    uint32_t bits = L.m;
    if (P.filter_mode != 0) {
      const PredF64 pr = pred_of(a);
      bits = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if ((uint32_t)k < L.cnt && ((L.m >> k) & 1) && pred_f64(pr, v[k])) bits |= 1u << k;
    }
    sel = bits;
  }
  if ((flags & kPfBits) && L.cnt) put_bits(a.out_bits, L.g, sel, (L.g & 7) == 0 && L.cnt == 8);
  if (!agg) return;  // wave-uniform
  const uint32_t use = (flags & kPfAggSel) ? sel : L.m;
  if (agg == 1) {
    double hi = 0.0, lo = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if ((use >> k) & 1) dd_add(hi, lo, v[k], 0.0);
    acc.n += (uint64_t)__builtin_popcount(use);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double oh = __shfl_xor(hi, o, kWave), ol = __shfl_xor(lo, o, kWave);
      if ((lane_id() & o) == 0) {
        dd_add(hi, lo, oh, ol);
      } else {
        double h2 = oh, l2 = ol;
        dd_add(h2, l2, hi, lo);
        hi = h2;
        lo = l2;
      }
    }
    if (lane_id() == 0) {
      double *pp = a.partials + 2ull * (c.rtile0 + t);
      pp[0] = hi;
      pp[1] = lo;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (((use >> k) & 1) && !(v[k] != v[k])) {
        acc.mn = min(acc.mn, f64_key(v[k]));
        acc.mx = max(acc.mx, f64_key(v[k]));
      }
  }
}

// The span's aggregates into the chunk's accumulators (one atomic each).
template <int KIND>
__device__ __forceinline__ void span_flush(const PlanDesc &P, uint32_t chunk, SpanAcc &acc) {
  const uint32_t flags = P.flags;
  const int agg = (int)((flags >> kPfAggShift) & 3u);
  if (!agg) return;
  uint64_t *ak = P.a.acc + 4ull * chunk;
  if (agg == 1) {
    constexpr bool f64 = KIND == kSpanValues || KIND == kSpanRle;
    if (f64 && !(flags & kPfAggSel)) return;  // the partials hold the sums
    const uint64_t n = wave_sum_u64(acc.n);
    const uint64_t sum = f64 ? 0 : wave_sum_u64(acc.sum);
    if (lane_id() == 0 && n) {
      if (!f64) atomicAdd((unsigned long long *)ak, (unsigned long long)sum);
      atomicAdd((unsigned long long *)ak + 1, (unsigned long long)n);
    }
    return;
  }
  uint64_t mn = acc.mn, mx = acc.mx;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    mn = min(mn, (uint64_t)__shfl_xor((unsigned long long)mn, o, kWave));
    mx = max(mx, (uint64_t)__shfl_xor((unsigned long long)mx, o, kWave));
  }
  if (lane_id() == 0) {
    if (agg == 2) atomicMin((unsigned long long *)ak + 2, (unsigned long long)mn);
    else atomicMax((unsigned long long *)ak + 3, (unsigned long long)mx);
  }
}

// BUFF raw words of one lane's 8 records: NP plane words and, with TAIL, the
// two aligned u64 words holding its r-bit tail fields.  Every load is
// issued unconditionally (lanes past the chunk's end read its first group),
// so a tile's loads are in flight while the wave works on the previous one.
template <uint32_t NP>
struct BuffRaw {
  uint64_t w[NP ? NP : 1];
  uint64_t lo, hi;
};
template <uint32_t NP, bool TAIL>
__device__ __forceinline__ void buff_load(const DevChunk &c, uint32_t i0, BuffRaw<NP> &r) {
  const uint32_t i = i0 < c.n ? i0 : 0u;
  const EBI_GLOBAL uint8_t *base = as_global(reinterpret_cast<const uint8_t *>(c.data));
#pragma unroll
  for (uint32_t p = 0; p < NP; ++p)
    r.w[p] = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(base + (size_t)p * c.plane_stride + i);
  if (TAIL) {
    const EBI_GLOBAL uint8_t *tail = base + (size_t)NP * c.plane_stride;
    const uint32_t a = ((i / 8) * c.r) & ~7u;
    r.lo = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(tail + a);
    r.hi = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(tail + a + 8);
  }
}
// The deltas of buff_deltas from the raw words, as 32-bit values (NP <= 2:
// B < 24); zeros past the chunk's end.
template <uint32_t NP, bool TAIL>
__device__ __forceinline__ void buff_assemble32(const DevChunk &c, uint32_t i0, const BuffRaw<NP> &r,
                                                uint32_t d[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) d[j] = 0;
  if (i0 >= c.n) return;
#pragma unroll
  for (uint32_t p = 0; p < NP; ++p) {
    const uint64_t w = r.w[p] ^ 0x8080808080808080ull;  // planes store each byte XOR 0x80
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] = (d[j] << 8) | (uint32_t)((w >> (8 * j)) & 0xff);
  }
  if (TAIL) {
    const uint32_t sh = (((i0 / 8) * c.r) & 7u) * 8;
    const uint64_t bits = sh ? (r.lo >> sh) | (r.hi << (64 - sh)) : r.lo;
    const uint32_t mk = (1u << c.r) - 1;
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] = (d[j] << c.r) | ((uint32_t)(bits >> (j * c.r)) & mk);
  }
}

template <uint32_t NP, bool TAIL>
__device__ __forceinline__ void span_buff(const PlanDesc &P, const DevChunk &c, const CodeSel &cs, uint32_t t0,
                                          uint32_t nt, SpanAcc &acc) {
  const uint32_t lane = lane_id();
  BuffRaw<NP> cur, nxt;
  buff_load<NP, TAIL>(c, t0 * kTile + lane * 8, cur);
  for (uint32_t k = 0; k < nt; ++k) {
    const uint32_t t = t0 + k, tn = k + 1 < nt ? t + 1 : t;  // the last tile reloads itself
    buff_load<NP, TAIL>(c, tn * kTile + lane * 8, nxt);
    uint32_t d[8];
    buff_assemble32<NP, TAIL>(c, t * kTile + lane * 8, cur, d);
    code_tile<kSpanBuff>(P, cs, tile_lane(P, c, t), d, acc);
    cur = nxt;
  }
}

// DeltaSprintz with W <= 16: each lane's 8 fields (W bytes from byte
// (i0 / 8) * W, MSB first) lie in three aligned u64 words; the words of tile
// k + 1 load during tile k.
struct SprRaw {
  uint64_t w0, w1, w2;
};
__device__ __forceinline__ void sprintz_load16(const DevChunk &c, uint32_t i0, SprRaw &r) {
  const uint32_t i = i0 < c.n ? i0 : 0u;
  const EBI_GLOBAL uint64_t *p =
      as_global(reinterpret_cast<const uint64_t *>(c.data)) + (((i / 8) * c.bits) >> 3);
  r.w0 = p[0];
  r.w1 = p[1];
  r.w2 = p[2];
}
// The 8 fields of this lane for a fixed W <= 16 as exact 32-bit inclusive
// prefix sums of their zigzag-decoded deltas: after the byte-aligned funnel
// shift every field sits at a fixed bit of four 32-bit words.  Fields past
// the chunk's end decode padding; they only feed the prefixes of records
// that are not selected (later fields of the last lane, empty lanes).
template <uint32_t W>
__device__ __forceinline__ void sprintz_local(uint32_t i0, const SprRaw &r, int32_t s[8]) {
  const uint32_t sh = (((i0 / 8) * W) & 7u) * 8;
  const uint64_t b0 = __builtin_bswap64(r.w0), b1 = __builtin_bswap64(r.w1), b2 = __builtin_bswap64(r.w2);
  const uint64_t hi = sh ? (b0 << sh) | (b1 >> (64 - sh)) : b0;  // the group's 8W bits, MSB first
  const uint64_t lo = sh ? (b1 << sh) | (b2 >> (64 - sh)) : b1;
  const uint32_t w[4] = {(uint32_t)(hi >> 32), (uint32_t)hi, (uint32_t)(lo >> 32), (uint32_t)lo};
  constexpr uint32_t mk = (1u << W) - 1;
  int32_t a = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t b = j * W, k = b >> 5, o = b & 31;
    const uint32_t z = o + W <= 32 ? (w[k] >> (32 - o - W)) & mk
                                   : ((w[k] << (o + W - 32)) | (w[k + 1 < 4 ? k + 1 : 3] >> (64 - o - W))) & mk;
    a += (int32_t)(z >> 1) ^ -(int32_t)(z & 1);
    s[j] = a;
  }
}

// DeltaSprintz with W > 16: field j of the lane's records starts at bit
// (i0 + j) * W (MSB first) and lies in the two aligned u64 words from its
// first bit's word; the loads of four fields are issued before they are cut
// out.
__device__ __forceinline__ void sprintz_q_wide(const DevChunk &c, uint32_t i0, int64_t tile_base, int64_t q[8]) {
  const uint32_t W = c.bits;
  const uint32_t cnt = i0 < c.n ? min(8u, c.n - i0) : 0u;
  const uint64_t b0 = (uint64_t)(cnt ? i0 : 0u) * W;
  const EBI_GLOBAL uint64_t *p = as_global(reinterpret_cast<const uint64_t *>(c.data));
  int64_t s = 0;
#pragma unroll
  for (int h = 0; h < 8; h += 4) {
    uint64_t w[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint64_t wi = (b0 + (uint64_t)(h + j) * W) >> 6;
      w[2 * j] = p[wi];
      w[2 * j + 1] = p[wi + 1];
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t o = (uint32_t)((b0 + (uint64_t)(h + j) * W) & 63);
      const uint64_t hi = __builtin_bswap64(w[2 * j]), lo = __builtin_bswap64(w[2 * j + 1]);
      const uint64_t x = o ? (hi << o) | (lo >> (64 - o)) : hi;
      const uint64_t z = x >> (64 - W);  // W <= 64: a shift of 0 .. 47
      const int64_t dj = (uint32_t)(h + j) < cnt ? (int64_t)(z >> 1) ^ -(int64_t)(z & 1) : 0;
      s = (int64_t)((uint64_t)s + (uint64_t)dj);
      q[h + j] = s;
    }
  }
  const int64_t incl = wave_incl_scan_i64(s);
  const int64_t excl = (int64_t)((uint64_t)incl - (uint64_t)s) + tile_base;
#pragma unroll
  for (int j = 0; j < 8; ++j) q[j] = (int64_t)((uint64_t)q[j] + (uint64_t)excl);
}

template <uint32_t W>
__device__ __forceinline__ void span_sprintz_w(const PlanDesc &P, const DevChunk &c, const CodeSel &cs,
                                               uint32_t t0, uint32_t nt, SpanAcc &acc) {
  const uint32_t lane = lane_id();
  const int64_t *tidx = P.view->tile_index + c.tile0;
  SprRaw cur, nxt;
  sprintz_load16(c, t0 * kTile + lane * 8, cur);
  for (uint32_t k = 0; k < nt; ++k) {
    const uint32_t t = t0 + k, tn = k + 1 < nt ? t + 1 : t;  // the last tile reloads itself
    sprintz_load16(c, tn * kTile + lane * 8, nxt);
    int32_t s[8];
    sprintz_local<W>(t * kTile + lane * 8, cur, s);
    const int32_t excl = (int32_t)wave_incl_scan_u32((uint32_t)s[7]) - s[7];  // every lane
    const int64_t base = (int64_t)((uint64_t)tidx[t] + (uint64_t)(int64_t)excl);
    sprintz_tile(P, cs, tile_lane(P, c, t), base, s, acc);
    cur = nxt;
  }
}

__device__ __forceinline__ void span_sprintz(const PlanDesc &P, const DevChunk &c, const CodeSel &cs, uint32_t t0,
                                             uint32_t nt, SpanAcc &acc) {
  const uint32_t lane = lane_id();
  const int64_t *tidx = P.view->tile_index + c.tile0;
  switch (c.bits) {
#define EBI_SPAN_SPR(w) \
  case w: span_sprintz_w<w>(P, c, cs, t0, nt, acc); return;
    EBI_SPAN_SPR(1) EBI_SPAN_SPR(2) EBI_SPAN_SPR(3) EBI_SPAN_SPR(4)
    EBI_SPAN_SPR(5) EBI_SPAN_SPR(6) EBI_SPAN_SPR(7) EBI_SPAN_SPR(8)
    EBI_SPAN_SPR(9) EBI_SPAN_SPR(10) EBI_SPAN_SPR(11) EBI_SPAN_SPR(12)
    EBI_SPAN_SPR(13) EBI_SPAN_SPR(14) EBI_SPAN_SPR(15) EBI_SPAN_SPR(16)
#undef EBI_SPAN_SPR
    default: break;
  }
  {  // W = 0 or W > 16
    for (uint32_t t = t0; t < t0 + nt; ++t) {
      int64_t q[8];
      if (c.bits == 0) {  // every delta 0: q = the tile's starting prefix
#pragma unroll
        for (int j = 0; j < 8; ++j) q[j] = tidx[t];
      } else {
        sprintz_q_wide(c, t * kTile + lane * 8, tidx[t], q);  // every lane (wave scan)
      }
      code_tile<kSpanSprintz>(P, cs, tile_lane(P, c, t), reinterpret_cast<const uint64_t *>(q), acc);
    }
  }
}

// Decoded records of one lane's 8 records (addresses clamped into the chunk,
// so every load is issued; values past the chunk's end are not used).  An
// Uncompressed chunk's body is its records: read in place.
__device__ __forceinline__ void values_load(const DevChunk &c, const double *values, uint32_t i0, double v[8]) {
  const uint32_t cnt = i0 < c.n ? min(8u, c.n - i0) : 0u;
  const double *src = (c.codec == kCodecPlain ? reinterpret_cast<const double *>(c.data) : values + c.rec_off) +
                      (cnt ? i0 : 0u);
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) v[k] = src[cnt ? min(k, cnt - 1) : 0u];
}

__device__ __forceinline__ void span_values(const PlanDesc &P, const DevChunk &c, uint32_t t0, uint32_t nt,
                                            SpanAcc &acc) {
  const uint32_t lane = lane_id();
  const uint32_t flags = P.flags;
  const bool load = ((flags & kPfPred) && P.filter_mode != 0) || ((flags >> kPfAggShift) & 3u);
  if (!load) {
    for (uint32_t t = t0; t < t0 + nt; ++t) values_tile(P, c, t, nullptr, acc);
    return;
  }
  double cur[8], nxt[8];
  values_load(c, P.a.values, t0 * kTile + lane * 8, cur);
  for (uint32_t k = 0; k < nt; ++k) {
    const uint32_t t = t0 + k, tn = k + 1 < nt ? t + 1 : t;
    values_load(c, P.a.values, tn * kTile + lane * 8, nxt);
    values_tile(P, c, t, cur, acc);
#pragma unroll
    for (int j = 0; j < 8; ++j) cur[j] = nxt[j];
  }
}

// Bits of global bitset word w inside records [g, g + r).
__device__ __forceinline__ uint64_t range_word(uint64_t w, uint64_t g, uint32_t r) {
  const uint64_t lo = w << 6, a = max(g, lo), b = min(g + r, lo + 64);
  if (a >= b) return 0;
  const uint32_t n = (uint32_t)(b - a);
  return (n == 64 ? ~0ull : (1ull << n) - 1) << (uint32_t)(a - lo);
}

// RLE queried per run (chunk_reader/run_length.rs:68-91 expands runs; the
// Decoder's default filter / sum / min / max then see each record of a run):
// a run's value is tested once, its records' bits are set a word at a time
// (ANDed with the bitmask), and the run adds value x count (an exact
// two-product into the tile's double-double partial) or its key once.
// Field tiles: lane l of tile t owns fields t*kTile + 8l .. +7, whose first
// record follows from the tile's start (tile index) and a wave scan of the
// run lengths.  A lane's 8 run bytes and 8 values are read as one u64 and
// four 16-byte loads, all issued before the first is used.
struct RleRaw {
  uint64_t runs;
  double2 v[4];
};
__device__ __forceinline__ void rle_load(const DevChunk &c, uint32_t f0, RleRaw &r) {
  const uint32_t f = f0 < c.fields ? f0 : 0u;  // f0 is a multiple of 8: 8-byte runs, 64-byte values
  const EBI_GLOBAL uint8_t *base = as_global(reinterpret_cast<const uint8_t *>(c.data));
  r.runs = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(base + f);
  const EBI_GLOBAL double2 *v = reinterpret_cast<const EBI_GLOBAL double2 *>(base + c.values_off) + f / 2;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.v[k] = v[k];
}

__device__ __forceinline__ void span_rle(const PlanDesc &P, const DevChunk &c, uint32_t t0, uint32_t nt,
                                         SpanAcc &acc) {
  const QueryArgs &a = P.a;
  const uint32_t lane = lane_id(), flags = P.flags;
  const int agg = (int)((flags >> kPfAggShift) & 3u);
  const bool pred = (flags & kPfPred) && P.filter_mode != 0;
  const bool bits_out = (flags & kPfBits) != 0;
  const PredF64 pr = pred_of(a);
  const int64_t *tidx = P.view->tile_index + c.tile0;
  for (uint32_t t = t0; t < t0 + nt; ++t) {
    const uint32_t f0 = t * kTile + lane * 8;
    RleRaw cur;
    rle_load(c, f0, cur);
    uint32_t run[8];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // 0 past the chunk's fields
      run[j] = f0 + j < c.fields ? (uint32_t)(cur.runs >> (8 * j)) & 0xffu : 0u;
      s += run[j];
    }
    const double v8[8] = {cur.v[0].x, cur.v[0].y, cur.v[1].x, cur.v[1].y,
                          cur.v[2].x, cur.v[2].y, cur.v[3].x, cur.v[3].y};
    const uint32_t incl = wave_incl_scan_u32(s);
    uint32_t pos = (uint32_t)tidx[t] + (incl - s);
    double hi = 0.0, lo = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t r = pos < c.n ? min(run[j], c.n - pos) : 0u;  // run totals are checked at load
      if (r) {
        const double v = v8[j];
        if (!pred || pred_f64(pr, v)) {
          const uint64_t g = c.rec_off + pos;
          uint32_t cnt = r;
          if (a.mask || bits_out) {
            cnt = 0;
            for (uint64_t w = g >> 6; w <= (g + r - 1) >> 6; ++w) {
              uint64_t b = range_word(w, g, r);
              if (a.mask) b &= a.mask[w];
              cnt += (uint32_t)__popcll(b);
              if (bits_out && b) atomicOr(reinterpret_cast<unsigned long long *>(a.out_bits + w), b);
            }
          }
          if (agg == 1) {
            if (cnt) {
              const double p = v * (double)cnt, e = fma(v, (double)cnt, -p);  // v x cnt exactly
              dd_add(hi, lo, p, e);
            }
            acc.n += cnt;
          } else if (agg && cnt && !(v != v)) {  // Rust f64::min / max ignore NaN
            if (agg == 2) acc.mn = min(acc.mn, f64_key(v));
            else acc.mx = max(acc.mx, f64_key(v));
          }
        }
      }
      pos += run[j];
    }
    if (agg == 1) {  // wave-uniform
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
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
      if (lane == 0) {
        double *pp = a.partials + 2ull * (c.rtile0 + t);
        pp[0] = hi;
        pp[1] = lo;
      }
    }
  }
}

// The record-tile pass of the plans of one SpanKind (their span jobs
// concatenated): one wave per (plan, span job).  One kernel per kind, so
// each gets the registers its own path needs.
template <int KIND>
__global__ void __launch_bounds__(256) k_group_tiles(const PlanDesc *__restrict__ plans,
                                                    const uint32_t *__restrict__ prefix, uint32_t n_plans,
                                                    uint32_t n_jobs) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_jobs) return;
  const uint32_t p = plan_of(prefix, n_plans, j);
  const PlanDesc &P = plans[p];
  const Job jb = P.span_jobs[j - prefix[p]];
  if (!P.a.chunk_sel[jb.chunk]) return;  // wave-uniform
  const DevChunk c = P.view->chunks[jb.chunk];
  SpanAcc acc;
  if (KIND == kSpanValues || KIND == kSpanRle) {
    if (KIND == kSpanValues) span_values(P, c, jb.tile, jb.kind, acc);
    else span_rle(P, c, jb.tile, jb.kind, acc);
    span_flush<KIND>(P, jb.chunk, acc);
    return;
  }
  ChunkPred cp{};
  if ((P.flags & kPfPred) && P.filter_mode == 1) cp = P.a.cpred[jb.chunk];
  const CodeSel cs = KIND == kSpanBuff ? code_sel_buff(P, cp) : code_sel_sprintz(P, cp);
  if (cs.none) return;  // no record: the bitset is zeroed, no aggregate
  if (KIND == kSpanSprintz) {
    span_sprintz(P, c, cs, jb.tile, jb.kind, acc);
  } else {
    switch (c.n_planes * 2 + (c.r ? 1u : 0u)) {
#define EBI_SPAN_BUFF(np)                                                    \
  case 2 * np: span_buff<np, false>(P, c, cs, jb.tile, jb.kind, acc); break; \
  case 2 * np + 1: span_buff<np, true>(P, c, cs, jb.tile, jb.kind, acc); break;
      EBI_SPAN_BUFF(0)
      EBI_SPAN_BUFF(1)
      EBI_SPAN_BUFF(2)
#undef EBI_SPAN_BUFF
      default:  // B >= 24: 64-bit deltas, a plane loop per tile, one tile in flight
        for (uint32_t t = jb.tile; t < jb.tile + jb.kind; ++t) {
          uint64_t d[8];
          buff_deltas(c, t * kTile + lane_id() * 8, d);
          code_tile<kSpanBuff>(P, cs, tile_lane(P, c, t), d, acc);
        }
        break;
    }
  }
  span_flush<KIND>(P, jb.chunk, acc);
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
// BUFF chunk sums are integers times 2^-frac.  While every integer and the
// sum of their magnitudes stay below 2^53, each partial sum of the Decoder's
// `result += v` is exact, so the in-order sum equals the integer total times
// 2^-frac and can be computed in any order.
constexpr uint64_t kExactLimit = 1ull << 53;
constexpr int64_t kNotExact = INT64_MIN;

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
      const uint32_t nt = (P.flags & kPfFieldTiles) ? (uint32_t)((c.fields + kTile - 1) / kTile)
                                                    : (c.n + kTile - 1) / kTile;
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
  int64_t exact = kNotExact;  // BUFF sums: the chunk's value is exact * 2^-frac
  if (take) {
    const DevChunk c = P.view->chunks[ci];
    const bool mask = (P.flags & (kPfMask | kPfAggSel)) != 0;
    if (c.codec == kCodecBUFF) {
      const double scale = ldexp(1.0, (int)c.frac);  // 2.0f64.powi(frac)
      if (agg == 1) {
        if (c.bits == 0) {  // buff/sum.rs:80-84: value * n, bitmask ignored
          v = ((double)c.base / scale) * (double)c.n;
          const uint64_t ab = c.base < 0 ? 0ull - (uint64_t)c.base : (uint64_t)c.base;
          if (ab < kExactLimit && (c.n == 0 || ab <= (kExactLimit - 1) / c.n)) exact = c.base * (int64_t)c.n;
        } else {
          const int64_t s = (int64_t)(ak[0] + (uint64_t)c.base * ak[1]);
          v = (double)s / scale;
          if (s > -(int64_t)kExactLimit && s < (int64_t)kExactLimit) exact = s;
        }
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
  cs[1] = as_f64((uint64_t)exact);
  const_cast<uint32_t *>(P.a.chunk_sel)[ci] = take;
}

// The Decoder's fold (api/decoder.rs:323-439) of every plan with kPfFold,
// one block per plan.  The block stages the chunks' results in LDS (all
// loads in flight at once), then:
//  * BUFF / DeltaSprintz sums in the reference's order: `result += v` chunk
//    after chunk, by one lane reading LDS 16 values ahead of its adds (an
//    untaken chunk is staged as -0.0, which leaves every sum as it is);
//  * f64 sums: lane l of wave 0 folds the double-double results of chunks
//    l, l + 64, ... then a butterfly in a lane-independent order
//    (deterministic);
//  * min / max: f64::min / f64::max from +-INFINITY, in any order.
constexpr uint32_t kFinStage = 2048;  // chunks per LDS stage (4096 for the in-order sums)
__global__ void __launch_bounds__(256) k_group_finish(const PlanDesc *plans) {
  __shared__ double2 buf[kFinStage];
  __shared__ uint8_t tk[kFinStage];
  const PlanDesc &P = plans[blockIdx.x];
  if (!(P.flags & kPfFold)) return;  // block-uniform
  const int agg = (int)((P.flags >> kPfAggShift) & 3u);
  const uint32_t tid = threadIdx.x, lane = lane_id();
  const uint32_t nc = P.n_chunks;
  const double *cs = P.chunk_sum;
  const uint32_t *take = P.a.chunk_sel;
  if (agg == 1 && !(P.flags & kPfF64Sum) && P.scheme == kCodecBUFF) {
    // the exact integer total when it applies (one frac, magnitudes < 2^53)
    __shared__ uint64_t r_sum[4], r_abs[4];
    __shared__ uint32_t r_fmin[4], r_fmax[4], r_bad[4];
    uint64_t sum = 0;  // wrapping; exact whenever the magnitudes stay below 2^53
    uint64_t mag = 0;
    uint32_t fmin = ~0u, fmax = 0, bad = 0;
    for (uint32_t k = tid; k < nc; k += blockDim.x) {
      if (!take[k]) continue;
      const int64_t x = (int64_t)as_u64(cs[2ull * k + 1]);
      if (x == kNotExact) {
        bad = 1;
        continue;
      }
      const uint32_t f = P.view->chunks[k].frac;
      fmin = min(fmin, f);
      fmax = max(fmax, f);
      sum += (uint64_t)x;
      mag += x < 0 ? 0ull - (uint64_t)x : (uint64_t)x;
      if (mag >= kExactLimit) {
        bad = 1;
        mag = kExactLimit;
      }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      sum += (uint64_t)__shfl_xor((unsigned long long)sum, o, kWave);
      mag = min(mag + (uint64_t)__shfl_xor((unsigned long long)mag, o, kWave), kExactLimit);
      fmin = min(fmin, (uint32_t)__shfl_xor((int)fmin, o, kWave));
      fmax = max(fmax, (uint32_t)__shfl_xor((int)fmax, o, kWave));
      bad |= (uint32_t)__shfl_xor((int)bad, o, kWave);
    }
    if (lane == 0) {
      const uint32_t w = tid / kWave;
      r_sum[w] = sum;
      r_abs[w] = mag;
      r_fmin[w] = fmin;
      r_fmax[w] = fmax;
      r_bad[w] = bad;
    }
    __syncthreads();
    if (tid == 0) {
      for (uint32_t w = 1; w < blockDim.x / kWave; ++w) {
        sum += r_sum[w];
        mag = min(mag + r_abs[w], kExactLimit);
        fmin = min(fmin, r_fmin[w]);
        fmax = max(fmax, r_fmax[w]);
        bad |= r_bad[w];
      }
      r_bad[0] = bad || mag >= kExactLimit || (fmin != ~0u && fmin != fmax);
      if (!r_bad[0]) P.result->scalar = fmin == ~0u ? 0.0 : ldexp((double)(int64_t)sum, -(int)fmin);
    }
    __syncthreads();
    if (!r_bad[0]) return;  // block-uniform
  }
  if (agg == 1 && !(P.flags & kPfF64Sum)) {
    double *seq = reinterpret_cast<double *>(buf);
    constexpr uint32_t cap = 2 * kFinStage;
    double r = 0.0;
    for (uint32_t k0 = 0; k0 < nc; k0 += cap) {
      const uint32_t m = min(cap, nc - k0), mp = (m + 15) & ~15u;
#pragma unroll 8
      for (uint32_t i = tid; i < mp; i += blockDim.x)
        seq[i] = i < m && take[k0 + i] ? cs[2ull * (k0 + i)] : -0.0;
      __syncthreads();
      if (tid == 0) {
        const double2 *s2 = reinterpret_cast<const double2 *>(seq);
        const uint32_t nb = mp / 16;
        double2 x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = s2[j];
        for (uint32_t b = 0; b < nb; ++b) {
          const uint32_t b1 = b + 1 < nb ? b + 1 : b;
          double2 y[8];
#pragma unroll
          for (int j = 0; j < 8; ++j) y[j] = s2[8 * b1 + j];
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            r += x[j].x;
            r += x[j].y;
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) x[j] = y[j];
        }
      }
      __syncthreads();
    }
    if (tid == 0) P.result->scalar = r;
    return;
  }
  const bool f64sum = (P.flags & kPfF64Sum) != 0;
  double hi = 0.0, lo = 0.0;
  double r = agg == 2 ? INFINITY : -INFINITY;
  for (uint32_t k0 = 0; k0 < nc; k0 += kFinStage) {
    const uint32_t m = min(kFinStage, nc - k0);
#pragma unroll 8
    for (uint32_t i = tid; i < m; i += blockDim.x) {
      tk[i] = take[k0 + i] != 0;
      buf[i] = make_double2(cs[2ull * (k0 + i)], cs[2ull * (k0 + i) + 1]);
    }
    __syncthreads();
    if (tid < kWave) {
      for (uint32_t i = lane; i < m; i += kWave) {
        if (!tk[i]) continue;
        if (f64sum) dd_add(hi, lo, buf[i].x, buf[i].y);
        else r = agg == 2 ? rmin_d(r, buf[i].x) : rmax_d(r, buf[i].x);
      }
    }
    __syncthreads();
  }
  if (tid >= kWave) return;
  double out;
  if (f64sum) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
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
    out = isfinite(hi) ? hi + lo : hi;
  } else {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      const double x = __shfl_xor(r, o, kWave);
      r = agg == 2 ? rmin_d(r, x) : rmax_d(r, x);
    }
    out = r;
  }
  if (lane == 0) P.result->scalar = out;
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

void group_prep(hipStream_t s, const GroupFill *fills, uint32_t n_fills, uint64_t max_words, const PlanDesc *plans,
                const uint32_t *prefix, uint32_t n_plans, uint32_t n_items) {
  const uint32_t fbx = n_fills ? (uint32_t)std::max<uint64_t>(std::min<uint64_t>((max_words + 255) / 256, 1024), 1) : 0;
  const uint64_t blocks = (uint64_t)n_fills * fbx + (n_items + 3) / 4;
  if (blocks)
    hipLaunchKernelGGL(k_group_prep, dim3((uint32_t)blocks), dim3(256), 0, s, fills, n_fills, std::max(fbx, 1u), plans,
                       prefix, n_plans, n_items);
}
void group_tiles(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_jobs,
                 int kind) {
  if (!n_jobs) return;
  const dim3 grid((n_jobs + 3) / 4), block(256);
  if (kind == kSpanBuff)
    hipLaunchKernelGGL(k_group_tiles<kSpanBuff>, grid, block, 0, s, plans, prefix, n_plans, n_jobs);
  else if (kind == kSpanSprintz)
    hipLaunchKernelGGL(k_group_tiles<kSpanSprintz>, grid, block, 0, s, plans, prefix, n_plans, n_jobs);
  else if (kind == kSpanRle)
    hipLaunchKernelGGL(k_group_tiles<kSpanRle>, grid, block, 0, s, plans, prefix, n_plans, n_jobs);
  else
    hipLaunchKernelGGL(k_group_tiles<kSpanValues>, grid, block, 0, s, plans, prefix, n_plans, n_jobs);
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
