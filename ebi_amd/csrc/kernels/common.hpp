// Device helpers shared by the gfx950 kernels of libebi_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../device.hpp"

namespace ebi {

constexpr uint64_t kNan = 0x7ff8000000000000ull;
constexpr int kWave = 64;

__device__ __forceinline__ double as_f64(uint64_t u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ uint64_t as_u64(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Chimp / Chimp128 / ElfXor 3-bit leading-zero codes (chunk_reader/chimp.rs:40).
__device__ __forceinline__ uint32_t lead_repr(uint32_t code) {
  // 0, 8, 12, 16, 18, 20, 22, 24 packed as bytes.
  const uint64_t t = 0x18161412100C0800ull;
  return (uint32_t)(t >> (8 * code)) & 0xff;
}

// MSB-first reader over one lane's own bit stream.  `base` is 4-byte aligned
// (chunk bodies are 256-byte aligned); reads may run up to 8 bytes past the
// current position, which the per-chunk zero padding absorbs.
struct LaneReader {
  const uint32_t *w;
  uint64_t buf;
  uint32_t nb;
  uint32_t pos;

  __device__ __forceinline__ void init(const uint8_t *base, uint32_t bitpos) {
    w = reinterpret_cast<const uint32_t *>(base) + (bitpos >> 5);
    buf = 0;
    nb = 0;
    pos = bitpos;
    fill();
    const uint32_t s = bitpos & 31;
    buf <<= s;
    nb -= s;
  }
  __device__ __forceinline__ void fill() {
    while (nb <= 32) {
      buf |= (uint64_t)__builtin_bswap32(*w++) << (32 - nb);
      nb += 32;
    }
  }
  // k in [0, 32]
  __device__ __forceinline__ uint32_t peek(uint32_t k) {
    fill();
    return k ? (uint32_t)(buf >> (64 - k)) : 0u;
  }
  // k in [0, 32], only after peek/fill
  __device__ __forceinline__ void skip(uint32_t k) {
    buf <<= k;
    nb -= k;
    pos += k;
  }
  __device__ __forceinline__ uint32_t read(uint32_t k) {
    fill();
    if (!k) return 0;
    const uint32_t v = (uint32_t)(buf >> (64 - k));
    buf <<= k;
    nb -= k;
    pos += k;
    return v;
  }
  // k in [0, 64]
  __device__ __forceinline__ uint64_t read64(uint32_t k) {
    if (k > 32) {
      const uint64_t hi = read(k - 32);
      return (hi << 32) | read(32);
    }
    return read(k);
  }
};

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (waves of one block run independent jobs, so no block barrier is used).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide inclusive scans (64 lanes).
__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t x) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int64_t y = __shfl_up(x, d, kWave);
    if ((int)lane_id() >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t y = __shfl_up(x, d, kWave);
    if ((int)lane_id() >= d) x += y;
  }
  return x;
}
__device__ __forceinline__ int64_t wave_sum_i64(int64_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
  return x;
}

}  // namespace ebi
