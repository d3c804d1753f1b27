// FP64 DGEMM on the gfx950 matrix cores for the decompress-then-matmul path.
//
// C_b = A_b * B_b for b in [0, batch), all row-major: A_b is m x k (lda = k),
// B_b is k x n (ldb = n), C_b is m x n (ldc = n); per-batch element strides
// sA, sB, sC (sB = 0 shares one B).  The reference computes the same product
// with cuBLAS after a CPU decode (cuda-binding/src/lib.rs:59-140, called from
// ebi/src/api/decoder.rs:659-724).
//
// Block tile 128 x 128 x 16, 256 threads = 4 waves in a 2 x 2 grid, each wave
// a 64 x 64 sub-tile of 4 x 4 v_mfma_f64_16x16x4f64 accumulators.  Operand
// tiles are staged through LDS (A transposed to k-major so a wave's fragment
// read is 16 consecutive doubles per k row), with the next K-slab prefetched
// into registers while the current one is multiplied.  Tile order is
// XCD-aware: workgroups are dispatched round-robin over the 8 XCDs, so the
// linear tile id is remapped to give each XCD a contiguous range of tiles,
// grouped 8 tile-rows high, and tiles sharing A rows / B columns meet in the
// same L2.
//
// v_mfma_f64_16x16x4f64 lane layout (wave64): A operand lane l holds
// A[l % 16][l / 16], B operand lane l holds B[l / 16][l % 16], and the four
// accumulator doubles of lane l hold D[4 * r + l / 16][l % 16], r = 0..3.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "launch.hpp"

namespace ebi {

typedef double v4f64 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128;
constexpr int kBN = 128;
constexpr int kBK = 16;
constexpr int kLdsA = kBM + 1;  // one pad double per k row: transposed stores spread over banks
constexpr int kLdsB = kBN + 1;
constexpr int kGroupM = 8;

__global__ void __launch_bounds__(256) k_dgemm(const double *__restrict__ A, const double *__restrict__ B,
                                               double *__restrict__ C, uint32_t m, uint32_t n, uint32_t k,
                                               uint64_t sA, uint64_t sB, uint64_t sC, uint32_t tiles_m,
                                               uint32_t tiles_n, uint32_t batch0) {
  __shared__ double As[kBK][kLdsA];
  __shared__ double Bs[kBK][kLdsB];

  const uint64_t b = (uint64_t)batch0 + blockIdx.y;
  A += b * sA;
  B += b * sB;
  C += b * sC;

  // XCD-aware tile order (see the file comment).
  const uint32_t nb = gridDim.x;
  uint32_t lid = blockIdx.x;
  if ((nb & 7u) == 0) lid = (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
  const uint32_t per_group = kGroupM * tiles_n;
  const uint32_t group = lid / per_group;
  const uint32_t first_m = group * kGroupM;
  const uint32_t gsize = min(tiles_m - first_m, (uint32_t)kGroupM);
  const uint32_t in_group = lid - group * per_group;
  const uint32_t tm = first_m + in_group % gsize;
  const uint32_t tn = in_group / gsize;
  const uint32_t m0 = tm * kBM, n0 = tn * kBN;

  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & 63u;
  const uint32_t wave = t >> 6;
  const uint32_t wm = (wave >> 1) * 64, wn = (wave & 1u) * 64;
  const uint32_t li = lane & 15u, lk = lane >> 4;

  // Global -> register staging: A slab 128 x 16 (8 doubles per thread, a
  // row's 16 k values are 16 consecutive threads), B slab 16 x 128 (8 per
  // thread, a row's 128 values are 128 consecutive threads).
  double ra[8], rb[8];
  auto load_slab = [&](uint32_t k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t idx = t + 256u * r;
      const uint32_t row = idx >> 4, kc = idx & 15u;
      const uint32_t gi = m0 + row, gk = k0 + kc;
      ra[r] = (gi < m && gk < k) ? A[(uint64_t)gi * k + gk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t idx = t + 256u * r;
      const uint32_t kr = idx >> 7, col = idx & 127u;
      const uint32_t gk = k0 + kr, gj = n0 + col;
      rb[r] = (gk < k && gj < n) ? B[(uint64_t)gk * n + gj] : 0.0;
    }
  };
  auto store_slab = [&]() {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t idx = t + 256u * r;
      As[idx & 15u][idx >> 4] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t idx = t + 256u * r;
      Bs[idx >> 7][idx & 127u] = rb[r];
    }
  };

  v4f64 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f64{0.0, 0.0, 0.0, 0.0};

  const uint32_t n_slabs = (k + kBK - 1) / kBK;
  load_slab(0);
  store_slab();
  __syncthreads();
  for (uint32_t s = 0; s < n_slabs; ++s) {
    if (s + 1 < n_slabs) load_slab((s + 1) * kBK);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      double a[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk + lk][wm + i * 16 + li];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Bs[kk + lk][wn + j * 16 + li];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    if (s + 1 < n_slabs) {
      store_slab();
      __syncthreads();
    }
  }

  // D[4 * r + lane / 16][lane % 16] of each 16 x 16 accumulator.
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t gj = n0 + wn + j * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t gi = m0 + wm + i * 16 + r * 4 + lk;
        if (gi < m && gj < n) C[(uint64_t)gi * n + gj] = acc[i][j][r];
      }
    }
  }
}

namespace launch {

hipError_t dgemm(hipStream_t s, const double *A, const double *B, double *C, uint32_t m, uint32_t n, uint32_t k,
                 uint64_t batch, uint64_t sA, uint64_t sB, uint64_t sC) {
  if (!m || !n || !batch) return hipSuccess;
  const uint32_t tiles_m = (m + kBM - 1) / kBM, tiles_n = (n + kBN - 1) / kBN;
  const uint64_t tiles = (uint64_t)tiles_m * tiles_n;
  if (tiles >= (1ull << 31)) return hipErrorInvalidValue;
  for (uint64_t b0 = 0; b0 < batch; b0 += 65535) {
    const uint32_t nbz = (uint32_t)((batch - b0) < 65535 ? (batch - b0) : 65535);
    hipLaunchKernelGGL(k_dgemm, dim3((uint32_t)tiles, nbz), dim3(256), 0, s, A, B, C, m, n, k, sA, sB, sC, tiles_m,
                       tiles_n, (uint32_t)b0);
  }
  return hipGetLastError();
}

}  // namespace launch
}  // namespace ebi
