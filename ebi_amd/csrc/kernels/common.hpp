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

// Decoded records are written once and not read back by the kernel: the
// tile decoders, whose every store instruction covers whole cache lines,
// stream them past the caches (non-temporal stores).  Partial-line writers
// (the XOR lane sinks, Chimp128) use ordinary stores, which L2 combines.
typedef double v2f64 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_pair(double *p, double a, double b) {
  __builtin_nontemporal_store(v2f64{a, b}, reinterpret_cast<v2f64 *>(p));
}
__device__ __forceinline__ void st_one(double *p, double a) { __builtin_nontemporal_store(a, p); }

// Chimp / Chimp128 / ElfXor 3-bit leading-zero codes (chunk_reader/chimp.rs:40).
__device__ __forceinline__ uint32_t lead_repr(uint32_t code) {
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
  const uint4 *p;  // next block to fetch
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
    p = reinterpret_cast<const uint4 *>(base) + (w / (4 * kBlk)) * kBlk;
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
// bit 0 of the window is the MSB of `hi`.  refill() tops the window up to at
// least 96 valid bits with 32-bit words from 16-byte block loads (one block
// requested ahead), so a record of up to 96 bits decodes from registers with
// shifts and selects, without per-field refills.  Reads at most 48 bytes past
// the current position.
struct BitWindow {
  const uint4 *p;
  uint64_t c0, c1;  // current 16-byte block as two big-endian halves
  uint4 nxt;
  uint32_t wi;      // next word of the current block (0 .. 4)
  uint64_t hi, lo;
  uint32_t nb;      // valid bits
  uint32_t pos;     // stream bit position of the window's first bit

  // Every step below is branch-free (selects and shifts with counts kept in
  // [0, 63]) except the block reload once per four words.  The reload
  // byte-swaps the prefetched block into c0, c1 (a computation, not a copy,
  // so the prefetch load can reuse nxt's registers and is waited for only at
  // the next reload).
  __host__ __device__ __forceinline__ void load_next() {
    c0 = __builtin_bswap64((uint64_t)nxt.x | ((uint64_t)nxt.y << 32));
    c1 = __builtin_bswap64((uint64_t)nxt.z | ((uint64_t)nxt.w << 32));
    nxt = *p++;
    wi = 0;
  }
  // append the next stream word when `need` (window bit nb, nb < 96)
  __host__ __device__ __forceinline__ void push_if(bool need) {
    const uint64_t half = (wi & 2) ? c1 : c0;
    const uint64_t w = (uint32_t)(half >> ((~wi & 1) * 32));
    const uint64_t top = w << 32;  // the word in the top 32 bits
    const uint64_t h = nb < 64 ? top >> (nb & 63) : 0;             // bits landing in hi
    const uint64_t l = nb > 32 ? w << ((96 - nb) & 63) : 0;        // bits landing in lo
    if (need) {
      hi |= h;
      lo |= l;
      nb += 32;
      ++wi;
      if (wi == 4) load_next();
    }
  }
  __host__ __device__ __forceinline__ void refill() {  // to >= 96 bits from >= 14
    push_if(nb < 96);
    push_if(nb < 96);
    push_if(nb < 96);
  }
  __host__ __device__ __forceinline__ void init(const uint8_t *base, uint32_t bitpos) {
    const uint32_t w = bitpos >> 5;
    p = reinterpret_cast<const uint4 *>(base) + (w >> 2);
    nxt = *p++;
    load_next();
    wi = w & 3;
    hi = lo = 0;
    nb = 0;
    pos = bitpos & ~31u;
    refill();
    refill();
    consume(bitpos & 31);
    refill();
  }
  // the `len` bits at window offset o (o < 64, 1 <= len <= 64)
  __host__ __device__ __forceinline__ uint64_t field(uint32_t o, uint32_t len) const {
    const uint64_t top = o ? (hi << (o & 63)) | (lo >> ((64 - o) & 63)) : hi;
    return top >> ((64 - len) & 63);
  }
  // drop k bits (k <= 96, k <= nb)
  __host__ __device__ __forceinline__ void consume(uint32_t k) {
    const uint64_t h_small = k ? (hi << (k & 63)) | (lo >> ((64 - k) & 63)) : hi;
    const uint64_t h_big = lo << ((k - 64) & 63);
    hi = k >= 64 ? h_big : h_small;
    lo = k >= 64 ? 0 : lo << (k & 63);
    nb -= k;
    pos += k;
  }
};

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
