// Device helpers shared by the gfx950 kernels of libebi_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../device.hpp"

namespace ebi {

constexpr uint64_t kNan = 0x7ff8000000000000ull;
constexpr int kWave = 64;

__host__ __device__ __forceinline__ double as_f64(uint64_t u) { return __builtin_bit_cast(double, u); }
__host__ __device__ __forceinline__ uint64_t as_u64(double d) { return __builtin_bit_cast(uint64_t, d); }
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Global-memory pointers for hot loops: pointers that arrive through
// descriptors are generic, which makes the compiler emit FLAT accesses (each
// waited for on both the memory and the LDS counter); casting them to the
// global address space gives global loads and stores.
#if defined(__HIP_DEVICE_COMPILE__)
#define EBI_GLOBAL __attribute__((address_space(1)))
#else
#define EBI_GLOBAL
#endif
template <class T>
__host__ __device__ __forceinline__ EBI_GLOBAL T *as_global(T *p) {
  return (EBI_GLOBAL T *)p;
}

// Decoded records are written once and not read back by the kernel: the
// tile decoders, whose every store instruction covers whole cache lines,
// stream them past the caches (non-temporal stores).  Partial-line writers
// (the XOR lane sinks, Chimp128) use ordinary stores, which L2 combines.
typedef double v2f64 __attribute__((ext_vector_type(2)));
// p: decoded-record output in HBM (stored through the global address space)
__device__ __forceinline__ void st_pair(double *p, double a, double b) {
  __builtin_nontemporal_store(v2f64{a, b}, reinterpret_cast<EBI_GLOBAL v2f64 *>(as_global(p)));
}
__device__ __forceinline__ void st_one(double *p, double a) { __builtin_nontemporal_store(a, as_global(p)); }

// Chimp / Chimp128 / ElfXor 3-bit leading-zero codes (chunk_reader/chimp.rs:40).
__host__ __device__ __forceinline__ uint32_t lead_repr(uint32_t code) {
  // 0, 8, 12, 16, 18, 20, 22, 24 packed as bytes.
  const uint64_t t = 0x18161412100C0800ull;
  return (uint32_t)(t >> (8 * code)) & 0xff;
}

// MSB-first reader over one lane's own bit stream.  `base` is 16-byte
// aligned (chunk bodies are 256-byte aligned).  The stream is fetched in
// blocks of kBlk x 16 bytes with one block of lookahead, so the next block
// is in flight while the current one is consumed.  kBlk = 1 suits lanes
// that read short, neighbouring pieces (a wave then covers one contiguous
// range); kBlk = 4 suits lanes walking long private streams (XOR codecs):
// each 64-byte block is fetched in one burst, so a line is not touched again
// after the caches may have dropped it.  Reads run at most 2 blocks + 8
// bytes past the current position, inside the per-chunk zero padding.
template <int kBlk>
struct LaneReaderT {
  const EBI_GLOBAL uint4 *p;  // next block to fetch
  uint4 cur[kBlk], nxt[kBlk];
  uint32_t wi;     // next word of `cur` (0 .. 4 * kBlk)
  uint64_t buf;
  uint32_t nb;
  uint32_t pos;

  __device__ __forceinline__ void load(uint4 (&dst)[kBlk]) {
#pragma unroll
    for (int k = 0; k < kBlk; ++k) dst[k] = p[k];
    p += kBlk;
  }
  __device__ __forceinline__ void init(const uint8_t *base, uint32_t bitpos) {
    const uint32_t w = bitpos >> 5;
    p = as_global(reinterpret_cast<const uint4 *>(base)) + (w / (4 * kBlk)) * kBlk;
    load(cur);
    load(nxt);
    wi = w % (4 * kBlk);
    buf = 0;
    nb = 0;
    pos = bitpos;
    fill();
    const uint32_t s = bitpos & 31;
    buf <<= s;
    nb -= s;
  }
  __device__ __forceinline__ uint32_t next_word() {
    if (wi == 4 * kBlk) {
#pragma unroll
      for (int k = 0; k < kBlk; ++k) cur[k] = nxt[k];
      load(nxt);
      wi = 0;
    }
    uint4 q = cur[0];
#pragma unroll
    for (int k = 1; k < kBlk; ++k)
      if ((wi >> 2) == (uint32_t)k) q = cur[k];
    const uint32_t c = wi & 3;
    const uint32_t v = c == 0 ? q.x : c == 1 ? q.y : c == 2 ? q.z : q.w;
    ++wi;
    return v;
  }
  __device__ __forceinline__ void fill() {
    while (nb <= 32) {
      buf |= (uint64_t)__builtin_bswap32(next_word()) << (32 - nb);
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
using LaneReader = LaneReaderT<1>;     // short neighbouring reads (DeltaSprintz tiles)

// MSB-first reader of a long private stream (one lane's XOR segment): the
// current 16-byte block plus kDepth - 1 blocks already requested, rotated by
// static register moves, so each lane keeps kDepth loads in flight and a
// refill waits on memory only when the stream outruns the whole queue.
// Reads at most kDepth * 16 bytes past the current position (chunk bodies
// carry 256 zero bytes of padding).
template <int kDepth>
struct PrefetchReader {
  const uint4 *p;        // next block to request
  uint4 cur;
  uint4 q[kDepth - 1];   // requested blocks, oldest first
  uint32_t wi;           // next word of `cur` (0 .. 4)
  uint64_t buf;
  uint32_t nb;
  uint32_t pos;

  __device__ __forceinline__ void init(const uint8_t *base, uint32_t bitpos) {
    const uint32_t w = bitpos >> 5;
    p = reinterpret_cast<const uint4 *>(base) + (w >> 2);
    cur = *p++;
#pragma unroll
    for (int k = 0; k < kDepth - 1; ++k) q[k] = *p++;
    wi = w & 3;
    buf = 0;
    nb = 0;
    pos = bitpos;
    fill();
    const uint32_t s = bitpos & 31;
    buf <<= s;
    nb -= s;
  }
  __device__ __forceinline__ uint32_t next_word() {
    if (wi == 4) {
      cur = q[0];
#pragma unroll
      for (int k = 0; k < kDepth - 2; ++k) q[k] = q[k + 1];
      q[kDepth - 2] = *p++;
      wi = 0;
    }
    const uint32_t v = wi == 0 ? cur.x : wi == 1 ? cur.y : wi == 2 ? cur.z : cur.w;
    ++wi;
    return v;
  }
  __device__ __forceinline__ void fill() {
    while (nb <= 32) {
      buf |= (uint64_t)__builtin_bswap32(next_word()) << (32 - nb);
      nb += 32;
    }
  }
  __device__ __forceinline__ uint32_t peek(uint32_t k) {
    fill();
    return k ? (uint32_t)(buf >> (64 - k)) : 0u;
  }
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
  __device__ __forceinline__ uint64_t read64(uint32_t k) {
    if (k > 32) {
      const uint64_t hi = read(k - 32);
      return (hi << 32) | read(32);
    }
    return read(k);
  }
};
// 128-bit window over an MSB-first stream for branch-free field extraction:
// bit 0 of the window is the MSB of `hi`; (hi, lo) always hold the 128
// stream bits from `pos` on.  The stream is held as the current and the next
// 16-byte block (big-endian 64-bit halves c0 c1 | n0 n1) and the bit offset
// b of pos in the current block; refill() rebuilds the window from them with
// two funnel shifts, and once b passes the current block the next becomes
// current and the block after it -- requested one block earlier -- is
// byte-swapped in while the following one is requested.  So a record of up
// to 128 bits decodes from registers with shifts and selects, and a lane
// waits on memory only for a load issued kDepth blocks earlier.  Reads at
// most 16 (kDepth + 2) bytes past the current position.  The lockstep
// segment kernels use kDepth 1 (many waves hide the latency); the lane-per-
// chunk passes, one long chain per lane with few waves, a deeper queue.
template <int kDepth>
struct BitWindowT {
  const EBI_GLOBAL uint4 *p;  // next block to request
  uint64_t c0, c1, n0, n1; // current and next block
  uint4 q[kDepth];         // requested blocks (after the next), oldest first
  uint32_t b;              // bit offset of pos in the current block, [0, 128)
  uint32_t stride;         // blocks from one stream block to the next
  uint64_t hi, lo;
  uint32_t pos;            // stream bit position of the window's first bit
  static constexpr uint32_t nb = 128;  // valid bits of the window

  __host__ __device__ static __forceinline__ uint64_t be64(uint32_t x, uint32_t y) {
    return __builtin_bswap64((uint64_t)x | ((uint64_t)y << 32));
  }
  // bits s .. s + 63 of the 128-bit value (x, y), s in [0, 63]
  __host__ __device__ static __forceinline__ uint64_t funnel(uint64_t x, uint64_t y, uint32_t s) {
    return s ? (x << s) | (y >> ((64 - s) & 63)) : x;
  }
  __host__ __device__ __forceinline__ void advance() {
    c0 = n0;
    c1 = n1;
    n0 = be64(q[0].x, q[0].y);
    n1 = be64(q[0].z, q[0].w);
#pragma unroll
    for (int k = 0; k < kDepth - 1; ++k) q[k] = q[k + 1];
    q[kDepth - 1] = *p;
    p += stride;
    b -= 128;
  }
  // after consume(): the window at the new position
  __host__ __device__ __forceinline__ void refill() {
    if (b >= 128) advance();
    const bool first = b < 64;
    const uint32_t s = b & 63;
    const uint64_t x0 = first ? c0 : c1, x1 = first ? c1 : n0, x2 = first ? n0 : n1;
    hi = funnel(x0, x1, s);
    lo = funnel(x1, x2, s);
  }
  __host__ __device__ __forceinline__ void init(const uint8_t *base, uint32_t bitpos) {
    init_at(reinterpret_cast<const uint4 *>(base) + (bitpos >> 7), bitpos, 1);
  }
  // blk: the stream's block holding bit `bitpos`; consecutive blocks of the
  // stream lie `step` blocks apart (1, or kIlvRowBlocks in an ilv)
  __host__ __device__ __forceinline__ void init_at(const uint4 *blk, uint32_t bitpos, uint32_t step) {
    stride = step;
    p = as_global(blk);
    const uint4 c = p[0], n = p[step];
#pragma unroll
    for (int k = 0; k < kDepth; ++k) q[k] = p[(2 + k) * step];
    p += (2 + kDepth) * step;
    c0 = be64(c.x, c.y);
    c1 = be64(c.z, c.w);
    n0 = be64(n.x, n.y);
    n1 = be64(n.z, n.w);
    b = bitpos & 127;
    pos = bitpos;
    refill();
  }
  // the `len` bits at window offset o (o < 64, 1 <= len <= 64, o + len <= 128)
  __host__ __device__ __forceinline__ uint64_t field(uint32_t o, uint32_t len) const {
    return funnel(hi, lo, o) >> ((64 - len) & 63);
  }
  // drop k bits (k <= 128); refill() before the next field
  __host__ __device__ __forceinline__ void consume(uint32_t k) {
    b += k;
    pos += k;
  }
};
using BitWindow = BitWindowT<1>;
constexpr int kSerialDepth = 8;  // blocks in flight per lane in the lane-per-chunk passes
using SerialWindow = BitWindowT<kSerialDepth>;

#ifndef EBI_XOR_DEPTH
#define EBI_XOR_DEPTH 2
#endif
using StreamReader = PrefetchReader<EBI_XOR_DEPTH>;  // long private streams (XOR codecs)

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (waves of one block run independent jobs, so no block barrier is used).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide inclusive scans (64 lanes).
// Inclusive wave prefix sums with DPP (register moves, no LDS): row shifts
// by 1, 2, 4, 8 scan each 16-lane row (lanes without a source read 0), then
// row_bcast:15 adds lane 15 of rows 0 / 2 into rows 1 / 3 and row_bcast:31
// adds lane 31 into rows 2 and 3.
template <int kCtrl, int kRowMask, bool kBound>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kCtrl, kRowMask, 0xf, kBound);
}
template <int kCtrl, int kRowMask, bool kBound>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t x) {
  const uint32_t lo = dpp_u32<kCtrl, kRowMask, kBound>((uint32_t)x);
  const uint32_t hi = dpp_u32<kCtrl, kRowMask, kBound>((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t x) {
  x += dpp_u32<kRowShr1, 0xf, true>(x);
  x += dpp_u32<kRowShr2, 0xf, true>(x);
  x += dpp_u32<kRowShr4, 0xf, true>(x);
  x += dpp_u32<kRowShr8, 0xf, true>(x);
  x += dpp_u32<kRowBcast15, 0xa, false>(x);
  x += dpp_u32<kRowBcast31, 0xc, false>(x);
  return x;
}
__device__ __forceinline__ int64_t wave_incl_scan_i64(int64_t v) {
  uint64_t x = (uint64_t)v;  // wrapping, as the reference's release build
  x += dpp_u64<kRowShr1, 0xf, true>(x);
  x += dpp_u64<kRowShr2, 0xf, true>(x);
  x += dpp_u64<kRowShr4, 0xf, true>(x);
  x += dpp_u64<kRowShr8, 0xf, true>(x);
  x += dpp_u64<kRowBcast15, 0xa, false>(x);
  x += dpp_u64<kRowBcast31, 0xc, false>(x);
  return (int64_t)x;
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
