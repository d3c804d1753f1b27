// Analytics kernels of libebi_hip (gfx950) over decoded columns in HBM:
// the reference's ML queries (Decoder::knn1, Decoder::matmul and the chunk
// readers' distance_squared / dot_product) and multi-column aggregations of
// arithmetic expressions (Aggregation / Expr).  Every floating-point fold
// that the reference computes as a sequential iterator sum runs in the same
// order in one lane, so those results are bit-identical; expression sums
// are double-double per tile (within 1e-12 relative of a compensated sum).
// Paths cited are relative to lemolatoon/ebi.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "launch.hpp"

namespace ebi {
namespace {

// Largest i in [0, n) with prefix[i] <= x (prefix ascending, prefix[0] = 0).
__device__ __forceinline__ uint32_t upper_index(const uint64_t *prefix, uint32_t n, uint64_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t key_of(double x) {
  const uint64_t b = as_u64(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// The reference's `iter.zip(..).map(..).sum()` over f64: a sequential fold
// (started from -0.0, as core's Sum for f64), each term rounded before it is
// added (no contraction; the library is built with -ffp-contract=off).
__device__ __forceinline__ double fold_dist(const double *v, const double *t, uint32_t count) {
  double s = -0.0;
  for (uint32_t i = 0; i < count; ++i) {
    const double d = v[i] - t[i];
    s += d * d;
  }
  return s;
}
__device__ __forceinline__ double fold_dot(const double *v, const double *t, uint32_t count) {
  double s = -0.0;
  for (uint32_t i = 0; i < count; ++i) s += v[i] * t[i];
  return s;
}

// Records of a zip of decompressed[min(len, off)..] with `want` target values
// (query.rs:157-192: the offset is clamped to the decoded length, the zip
// stops at the shorter side).
__device__ __forceinline__ uint32_t zip_count(uint32_t n_eff, uint32_t off, uint32_t want) {
  const uint32_t o = min(n_eff, off);
  return min(want, n_eff - o);
}

// knn1 candidates (api/decoder.rs:469-533): one lane per candidate, in the
// reference's order; d[g] its squared distance.  A candidate is either the
// completion of a vector begun at the end of the previous chunk (the first
// candidate of a chunk with `straddle`; compared with `<`) or a whole vector
// inside the chunk at offset j * L (compared with `<=`).  The reference
// restarts each chunk's whole vectors at offset 0 even after a completion.
__global__ void __launch_bounds__(256) k_knn_dist(const double *values, const KnnChunk *chunks, uint32_t nc,
                                                  const uint64_t *cand0, uint64_t n_cand, const double *target,
                                                  uint32_t L, double *d, unsigned long long *dmin) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t key = ~0ull;
  if (g < n_cand) {
    const uint32_t ci = upper_index(cand0, nc, g);
    const KnnChunk c = chunks[ci];
    const uint32_t j = (uint32_t)(g - cand0[ci]);
    double dist;
    if (c.straddle && j == 0) {
      // previous chunk's partial vector: target[0 .. m1), then this chunk's
      // first records against target[L - r ..] (r = L - m1), added as two sums
      const uint32_t n1 = zip_count(c.prev_n_eff, c.prev_off, c.m1);
      const double p1 = fold_dist(values + c.prev_rec_off + min(c.prev_n_eff, c.prev_off), target, n1);
      const uint32_t r = L - c.m1;
      const uint32_t n2 = zip_count(c.n_eff, 0, r);
      dist = p1 + fold_dist(values + c.rec_off, target + (L - r), n2);
    } else {
      const uint32_t off = (j - c.straddle) * L;
      dist = fold_dist(values + c.rec_off + min(c.n_eff, off), target, zip_count(c.n_eff, off, L));
    }
    d[g] = dist;
    if (dist == dist) key = key_of(dist);
  }
  // wave minimum, then one atomic per wave
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) key = min(key, (uint64_t)__shfl_xor((unsigned long long)key, o, kWave));
  if (lane_id() == 0 && key != ~0ull) atomicMin(dmin, (unsigned long long)key);
}

// The first candidate whose distance equals the minimum, and the last one
// compared with `<=` whose distance equals it.
__global__ void __launch_bounds__(256) k_knn_pick(const KnnChunk *chunks, uint32_t nc, const uint64_t *cand0,
                                                  uint64_t n_cand, const double *d, const unsigned long long *dmin,
                                                  unsigned long long *first, unsigned long long *last_le) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t f = ~0ull, l = 0;
  if (g < n_cand && *dmin != ~0ull && key_of(d[g]) == *dmin && d[g] == d[g]) {
    f = g;
    const uint32_t ci = upper_index(cand0, nc, g);
    const bool strict = chunks[ci].straddle && g == cand0[ci];
    if (!strict) l = g + 1;  // 0 = none
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    f = min(f, (uint64_t)__shfl_xor((unsigned long long)f, o, kWave));
    l = max(l, (uint64_t)__shfl_xor((unsigned long long)l, o, kWave));
  }
  if (lane_id() == 0) {
    if (f != ~0ull) atomicMin(first, (unsigned long long)f);
    if (l) atomicMax(last_le, (unsigned long long)l);
  }
}

// Decoder::matmul (api/decoder.rs:606-656) and the chunk readers'
// dot_product (query.rs:175-192): out[(b * rows + row) * cols + j] =
// sequential dot of the record run at `off_b + row * data_cols` (clamped and
// zipped as the reference does) with target column j (column-first, length
// k).  One lane per output element; batch b lies in the chunk whose batch
// prefix holds it, at offset (b - batch0) * stride.
__global__ void __launch_bounds__(256) k_dot(const double *values, const DotChunk *chunks, uint32_t nc,
                                             const uint64_t *batch0, uint64_t n_batch, uint32_t stride,
                                             uint32_t rows, uint32_t data_cols, const double *target, uint32_t k,
                                             uint32_t cols, double *out) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t per_batch = (uint64_t)rows * cols;
  if (g >= n_batch * per_batch) return;
  const uint64_t b = g / per_batch;
  const uint32_t rj = (uint32_t)(g - b * per_batch);
  const uint32_t row = rj / cols, j = rj % cols;
  const uint32_t ci = upper_index(batch0, nc, b);
  const DotChunk c = chunks[ci];
  const uint32_t off = c.off0 + (uint32_t)(b - batch0[ci]) * stride + row * data_cols;
  out[g] = fold_dot(values + c.rec_off + min(c.n_eff, off), target + (uint64_t)j * k, zip_count(c.n_eff, off, k));
}

__global__ void k_dist_one(const double *values, uint64_t rec_off, uint32_t n_eff, uint32_t off,
                           const double *target, uint32_t len, int dot, double *out) {
  if (threadIdx.x || blockIdx.x) return;
  const double *v = values + rec_off + min(n_eff, off);
  const uint32_t cnt = zip_count(n_eff, off, len);
  *out = dot ? fold_dot(v, target, cnt) : fold_dist(v, target, cnt);
}

// Selected records per chunk: popcount of the bitset over each chunk's
// records (one wave per chunk).
__global__ void __launch_bounds__(256) k_chunk_count(const DevChunk *chunks, uint32_t nc, const uint64_t *bits,
                                                     uint64_t *counts) {
  const uint32_t ci = blockIdx.x * 4 + threadIdx.x / kWave;
  if (ci >= nc) return;
  const DevChunk c = chunks[ci];
  const uint64_t a = c.rec_off, b = c.rec_off + c.n;
  uint64_t cnt = 0;
  for (uint64_t w = (a >> 6) + lane_id(); b > a && w <= ((b - 1) >> 6); w += kWave) {
    uint64_t x = bits[w];
    const uint64_t lo = w << 6;
    if (a > lo) x &= ~0ull << (a - lo);
    if (b < lo + 64) x &= (1ull << (b - lo)) - 1;
    cnt += __popcll(x);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, kWave);
  if (lane_id() == 0) counts[ci] = cnt;
}

__device__ __forceinline__ void dd_sum(double &hi, double &lo, double b) {
  const double s = hi + b;
  if (!isfinite(s)) {
    hi = s;
    lo = 0.0;
    return;
  }
  const double v = s - hi;
  const double e = (hi - (s - v)) + (b - v);
  const double t = e + lo;
  hi = s + t;
  lo = t - (hi - s);
}
__device__ __forceinline__ void dd_merge(double &hi, double &lo, double bh, double bl) {
  const double s = hi + bh;
  if (!isfinite(s)) {
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

// Expression records of one aggregation (Aggregation::compute, api/
// decoder.rs:835-917; Expr::compute_chunk, decoder/expr.rs:67-123): record k
// of chunk i takes, for every column node, the k-th selected record of that
// column's chunk i (cols[c] + col_off[c * nc + i] + k); the program is
// postfix.  One wave per tile of 512 records of one chunk (tiles[t] =
// {chunk, first record}); kind 3 sum: the tile's double-double sum to
// partials[t]; 1 max / 2 min: keys to *acc; 0: the values to out (at the
// chunk's output offset).
__global__ void __launch_bounds__(256) k_expr(ExprProgram prog, const double *const *cols, const uint64_t *col_off,
                                              uint32_t nc, const uint64_t *len, const uint64_t *out_off,
                                              const uint2 *tiles, uint32_t n_tiles, int kind, double *partials,
                                              unsigned long long *acc, double *out) {
  const uint32_t t = blockIdx.x * 4 + threadIdx.x / kWave;
  if (t >= n_tiles) return;  // wave-uniform
  const uint32_t ci = tiles[t].x, k0 = tiles[t].y;
  const uint32_t lane = lane_id();
  const uint64_t n = len[ci];
  double hi = 0.0, lo = 0.0;
  uint64_t kmin = ~0ull, kmax = 0;
  for (uint32_t s = 0; s < kTile / kWave; ++s) {
    const uint64_t k = (uint64_t)k0 + s * kWave + lane;
    if (k >= n) break;  // lanes past the chunk's records (no shuffles below)
    double st[kExprStack];
    int sp = 0;
    for (uint32_t p = 0; p < prog.n; ++p) {  // wave-uniform program
      const ExprNode e = prog.nodes[p];
      if (e.kind == 0) {
        st[sp++] = cols[e.index][col_off[(uint64_t)e.index * nc + ci] + k];
      } else if (e.kind == 1) {
        st[sp++] = e.literal;
      } else {
        const double b = st[--sp], a = st[sp - 1];
        st[sp - 1] = e.kind == 2 ? a + b : e.kind == 3 ? a - b : e.kind == 4 ? a * b : a / b;
      }
    }
    const double x = st[0];
    if (kind == 3) {
      dd_sum(hi, lo, x);
    } else if (kind == 0) {
      out[out_off[ci] + k] = x;
    } else if (x == x) {  // f64::max / f64::min skip NaN
      kmin = min(kmin, key_of(x));
      kmax = max(kmax, key_of(x));
    }
  }
  if (kind == 3) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {  // lane-independent order: every lane gets the same pair
      const double oh = __shfl_xor(hi, o, kWave), ol = __shfl_xor(lo, o, kWave);
      if ((lane & o) == 0) {
        dd_merge(hi, lo, oh, ol);
      } else {
        double h2 = oh, l2 = ol;
        dd_merge(h2, l2, hi, lo);
        hi = h2;
        lo = l2;
      }
    }
    if (lane == 0) {
      partials[2ull * t] = hi;
      partials[2ull * t + 1] = lo;
    }
  } else if (kind == 1 || kind == 2) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      kmin = min(kmin, (uint64_t)__shfl_xor((unsigned long long)kmin, o, kWave));
      kmax = max(kmax, (uint64_t)__shfl_xor((unsigned long long)kmax, o, kWave));
    }
    if (lane == 0) {
      if (kind == 1 && kmax) atomicMax(acc, (unsigned long long)kmax);
      if (kind == 2 && kmin != ~0ull) atomicMin(acc, (unsigned long long)kmin);
    }
  }
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

namespace launch {

void knn(hipStream_t s, const double *values, const KnnChunk *chunks, uint32_t nc, const uint64_t *cand0,
         uint64_t n_cand, const double *target, uint32_t L, double *d, unsigned long long *red) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, red, ~0ull, 2u);  // dmin, first
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, red + 2, 0ull, 1u);  // last_le
  if (!n_cand) return;
  const uint32_t blocks = (uint32_t)((n_cand + 255) / 256);
  hipLaunchKernelGGL(k_knn_dist, dim3(blocks), dim3(256), 0, s, values, chunks, nc, cand0, n_cand, target, L, d,
                     red);
  hipLaunchKernelGGL(k_knn_pick, dim3(blocks), dim3(256), 0, s, chunks, nc, cand0, n_cand, d,
                     (const unsigned long long *)red, red + 1, red + 2);
}

void dot(hipStream_t s, const double *values, const DotChunk *chunks, uint32_t nc, const uint64_t *batch0,
         uint64_t n_batch, uint32_t stride, uint32_t rows, uint32_t data_cols, const double *target, uint32_t k,
         uint32_t cols, double *out) {
  const uint64_t n = n_batch * rows * cols;
  if (!n) return;
  hipLaunchKernelGGL(k_dot, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, values, chunks, nc, batch0, n_batch,
                     stride, rows, data_cols, target, k, cols, out);
}

void dist_one(hipStream_t s, const double *values, uint64_t rec_off, uint32_t n_eff, uint32_t off,
              const double *target, uint32_t len, int dot, double *out) {
  hipLaunchKernelGGL(k_dist_one, dim3(1), dim3(64), 0, s, values, rec_off, n_eff, off, target, len, dot, out);
}

void chunk_count(hipStream_t s, const DevChunk *chunks, uint32_t nc, const uint64_t *bits, uint64_t *counts) {
  if (nc) hipLaunchKernelGGL(k_chunk_count, dim3((nc + 3) / 4), dim3(256), 0, s, chunks, nc, bits, counts);
}

void expr(hipStream_t s, const ExprProgram &prog, const double *const *cols, const uint64_t *col_off, uint32_t nc,
          const uint64_t *len, const uint64_t *out_off, const uint2 *tiles, uint32_t n_tiles, int kind,
          double *partials, unsigned long long *acc, double *out) {
  if (kind == 1 || kind == 2)
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, acc, kind == 1 ? 0ull : ~0ull, 1u);
  if (n_tiles)
    hipLaunchKernelGGL(k_expr, dim3((n_tiles + 3) / 4), dim3(256), 0, s, prog, cols, col_off, nc, len, out_off,
                       tiles, n_tiles, kind, partials, acc, out);
}

}  // namespace launch
}  // namespace ebi
