// GPU encoders of libebi_hip (gfx950) for BUFF and DeltaSprintz: the
// reference's write path (ChunkWriter::write, encoder.rs:317-405) for the
// two codecs whose chunk layout is a fixed-width packing, so that every
// record's bits have a known place once per-chunk statistics are known.
//
//  * k_enc_stats: one wave per tile of 2048 records of one chunk reduces
//    what fixes the chunk's layout -- BUFF: min and max fixed-point value
//    (compressor/buff.rs:222-262); DeltaSprintz: the OR of the zigzagged
//    deltas, whose bit length is W (compressor/sprintz.rs:151-175) -- into
//    per-chunk accumulators with one atomic per wave;
//  * the host turns them into each chunk's header fields, byte size and
//    offset in the file;
//  * k_enc_buff / k_enc_sprintz: one lane per 8 records writes their bytes:
//    BUFF planes (8 bytes per plane: byte = ((delta >> shift) & 0xff) ^ 0x80)
//    and the r-bit LSB-first tail (8 records x r bits = r whole bytes);
//    DeltaSprintz 8 x W MSB-first bits = W whole bytes after the 13-byte
//    header.  Lane 0 of each chunk also writes the chunk header.
// Records never depend on their neighbours' encoding (DeltaSprintz takes the
// previous record's quantized value directly from the input), so no scan is
// needed.  Output bytes are identical to the host encoder's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../host/format.hpp"
#include "common.hpp"
#include "launch.hpp"

namespace ebi {
namespace {

constexpr uint32_t kEncTile = 2048;

// Saturating f64 -> i64 with NaN -> 0 (Rust `as i64`).
__device__ __forceinline__ int64_t sat_i64_dev(double d) {
  if (d != d) return 0;
  if (d >= 9223372036854775808.0) return INT64_MAX;
  if (d < -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}

// (v * scale).round() as i64, compressor/sprintz.rs:160 (round half away).
__device__ __forceinline__ int64_t quantize(double v, double scale) { return sat_i64_dev(round(v * scale)); }

__device__ __forceinline__ uint64_t zigzag_dev(int64_t d) { return (uint64_t)((d << 1) ^ (d >> 63)); }

__global__ void __launch_bounds__(256) k_enc_stats(const double *v, uint64_t n, uint64_t per_chunk, int codec,
                                                   int32_t frac, double scale, uint32_t tiles_per_chunk,
                                                   uint64_t n_tiles, EncStat *stats) {
  const uint64_t t = (uint64_t)blockIdx.x * 4 + threadIdx.x / kWave;
  if (t >= n_tiles) return;  // wave-uniform
  const uint64_t c = t / tiles_per_chunk;
  const uint64_t c0 = c * per_chunk, c1 = min(n, c0 + per_chunk);
  const uint64_t a = c0 + (t - c * tiles_per_chunk) * kEncTile, b = min(c1, a + kEncTile);
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  uint64_t orz = 0;
  for (uint64_t i = a + lane_id(); i < b; i += kWave) {
    if (codec == kCodecBUFF) {
      const int64_t f = buff_fixed(v[i], frac);
      mn = min(mn, f);
      mx = max(mx, f);
    } else {
      const int64_t q = quantize(v[i], scale);
      const int64_t p = i == c0 ? q : quantize(v[i - 1], scale);  // the first delta is 0
      orz |= zigzag_dev((int64_t)((uint64_t)q - (uint64_t)p));
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    mn = min(mn, (int64_t)__shfl_xor((long long)mn, o, kWave));
    mx = max(mx, (int64_t)__shfl_xor((long long)mx, o, kWave));
    orz |= (uint64_t)__shfl_xor((unsigned long long)orz, o, kWave);
  }
  if (lane_id() == 0) {
    EncStat *s = stats + c;
    if (codec == kCodecBUFF) {
      atomicMin(reinterpret_cast<long long *>(&s->a), (long long)mn);
      atomicMax(reinterpret_cast<long long *>(&s->b), (long long)mx);
    } else {
      atomicOr(reinterpret_cast<unsigned long long *>(&s->b), (unsigned long long)orz);
    }
  }
}

__global__ void k_enc_stats_init(EncStat *stats, uint64_t nc, int codec) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  stats[c].a = codec == kCodecBUFF ? INT64_MAX : 0;
  stats[c].b = codec == kCodecBUFF ? INT64_MIN : 0;
}

__device__ __forceinline__ void put_le(uint8_t *p, uint64_t v, uint32_t bytes) {
  for (uint32_t k = 0; k < bytes; ++k) p[k] = (uint8_t)(v >> (8 * k));
}

// One lane per 8 records of the chunks described by `chunks` (chunk c holds
// records [c * per_chunk, ...), its body at file + chunks[c].off).
__global__ void __launch_bounds__(256) k_enc_buff(const double *v, uint64_t n, uint64_t per_chunk,
                                                  const EncChunk *chunks, uint32_t groups_per_chunk,
                                                  uint64_t n_groups, uint8_t *file) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  const uint64_t c = g / groups_per_chunk;
  const EncChunk ch = chunks[c];
  const uint64_t c0 = c * per_chunk, cn = min(n, c0 + per_chunk) - c0;
  const uint64_t j0 = (g - c * groups_per_chunk) * 8;  // first record in the chunk
  uint8_t *out = file + ch.off;
  if (j0 == 0) {  // header: base lo, base hi, n, B, frac (u32 LE each)
    put_le(out, (uint64_t)ch.base, 8);
    put_le(out + 8, cn, 4);
    put_le(out + 12, ch.bits, 4);
    put_le(out + 16, ch.frac, 4);
  }
  if (j0 >= cn || ch.bits == 0) return;
  const uint32_t cnt = (uint32_t)min<uint64_t>(8, cn - j0);
  uint64_t d[8];
#pragma unroll
  for (int k = 0; k < 8; ++k)
    d[k] = (uint32_t)k < cnt ? (uint64_t)buff_fixed(v[c0 + j0 + k], (int32_t)ch.frac) - (uint64_t)ch.base : 0;
  uint8_t *body = out + 20;
  const uint32_t planes = ch.bits >= 8 ? ch.bits / 8 : 0;
  const uint32_t r = ch.bits >= 8 ? ch.bits % 8 : ch.bits;
  for (uint32_t p = 0; p < planes; ++p) {
    const uint32_t sh = ch.bits - 8 * (p + 1);
    for (uint32_t k = 0; k < cnt; ++k) body[(uint64_t)p * cn + j0 + k] = (uint8_t)(((d[k] >> sh) & 0xff) ^ 0x80);
  }
  if (r) {  // 8 records x r bits, LSB first: r bytes (fewer for a short last group)
    uint64_t w = 0;
    for (uint32_t k = 0; k < cnt; ++k) w |= (d[k] & ((1ull << r) - 1)) << (k * r);
    uint8_t *tail = body + (uint64_t)planes * cn + (j0 / 8) * r;
    put_le(tail, w, (cnt * r + 7) / 8);
  }
}

__global__ void __launch_bounds__(256) k_enc_sprintz(const double *v, uint64_t n, uint64_t per_chunk,
                                                     const EncChunk *chunks, uint32_t groups_per_chunk,
                                                     uint64_t n_groups, double scale, uint64_t scale_bits,
                                                     uint8_t *file) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_groups) return;
  const uint64_t c = g / groups_per_chunk;
  const EncChunk ch = chunks[c];
  const uint64_t c0 = c * per_chunk, cn = min(n, c0 + per_chunk) - c0;
  const uint64_t j0 = (g - c * groups_per_chunk) * 8;
  uint8_t *out = file + ch.off;
  const uint32_t W = ch.bits;
  if (j0 == 0) {  // header, MSB-first: q0 (64 bits), scale (32 bits), W (8 bits)
    const uint64_t q0 = cn ? (uint64_t)quantize(v[c0], scale) : 0;
    for (int k = 0; k < 8; ++k) out[k] = (uint8_t)(q0 >> (56 - 8 * k));
    for (int k = 0; k < 4; ++k) out[8 + k] = (uint8_t)(scale_bits >> (24 - 8 * k));
    out[12] = (uint8_t)W;
  }
  if (j0 >= cn || W == 0) return;
  const uint32_t cnt = (uint32_t)min<uint64_t>(8, cn - j0);
  uint8_t *p = out + 13 + (j0 / 8) * W;  // 8 fields of W bits = W bytes
  int64_t prev = j0 == 0 ? quantize(v[c0], scale) : quantize(v[c0 + j0 - 1], scale);
  uint64_t acc = 0;  // pending bits, MSB-aligned count nb
  uint32_t nb = 0, o = 0;
  for (uint32_t k = 0; k < cnt; ++k) {
    const int64_t q = quantize(v[c0 + j0 + k], scale);
    const uint64_t z = zigzag_dev((int64_t)((uint64_t)q - (uint64_t)prev));
    prev = q;
    // append W bits of z, emitting whole bytes
    for (uint32_t left = W; left;) {
      const uint32_t take = min(left, 56u - nb);
      const uint64_t bits = (z >> (left - take)) & ((1ull << take) - 1);
      acc = (acc << take) | bits;
      nb += take;
      left -= take;
      while (nb >= 8) {
        p[o++] = (uint8_t)(acc >> (nb - 8));
        nb -= 8;
      }
      acc &= nb ? ((1ull << nb) - 1) : 0ull;
    }
  }
  if (nb) p[o] = (uint8_t)(acc << (8 - nb));  // zero padding of the chunk's last byte
}

}  // namespace

namespace launch {

void enc_stats(hipStream_t s, const double *v, uint64_t n, uint64_t per_chunk, uint64_t nc, int codec, int32_t frac,
               double scale, EncStat *stats) {
  if (nc) hipLaunchKernelGGL(k_enc_stats_init, dim3((uint32_t)((nc + 255) / 256)), dim3(256), 0, s, stats, nc, codec);
  const uint32_t tpc = (uint32_t)((per_chunk + kEncTile - 1) / kEncTile);
  const uint64_t n_tiles = nc * tpc;
  if (n_tiles)
    hipLaunchKernelGGL(k_enc_stats, dim3((uint32_t)((n_tiles + 3) / 4)), dim3(256), 0, s, v, n, per_chunk, codec,
                       frac, scale, tpc, n_tiles, stats);
}

void enc_pack(hipStream_t s, const double *v, uint64_t n, uint64_t per_chunk, uint64_t nc, int codec,
              const EncChunk *chunks, double scale, uint64_t scale_bits, uint8_t *file) {
  const uint32_t gpc = (uint32_t)((per_chunk + 7) / 8);
  const uint64_t n_groups = nc * std::max<uint64_t>(gpc, 1);
  if (!n_groups) return;
  const dim3 grid((uint32_t)((n_groups + 255) / 256));
  if (codec == kCodecBUFF)
    hipLaunchKernelGGL(k_enc_buff, grid, dim3(256), 0, s, v, n, per_chunk, chunks, std::max<uint32_t>(gpc, 1),
                       n_groups, file);
  else
    hipLaunchKernelGGL(k_enc_sprintz, grid, dim3(256), 0, s, v, n, per_chunk, chunks, std::max<uint32_t>(gpc, 1),
                       n_groups, scale, scale_bits, file);
}

}  // namespace launch
}  // namespace ebi
