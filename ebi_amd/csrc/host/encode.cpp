// Host write path of libebi_hip: the eight codec encoders and the EBI1 file
// writer.  It produces the files the GPU path decodes; it is not accelerated.
//
// Each encoder restates the reference compressor cited on it; chunks are
// encoded independently (every compressor resets per chunk, encoder.rs:156-157)
// so they run on a small host thread pool and are concatenated in order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/ebi_hip.h"
#include "format.hpp"

namespace ebi {
namespace {

inline unsigned clz64(uint64_t x) { return x ? (unsigned)__builtin_clzll(x) : 64u; }
inline unsigned ctz64(uint64_t x) { return x ? (unsigned)__builtin_ctzll(x) : 64u; }

// Saturating f64 -> i64 conversion with NaN -> 0 (Rust `as i64`).
inline int64_t sat_i64(double d) {
  if (std::isnan(d)) return 0;
  if (d >= 9223372036854775808.0) return INT64_MAX;
  if (d < -9223372036854775808.0) return INT64_MIN;
  return (int64_t)d;
}
inline uint32_t sat_u32(double d) {
  if (std::isnan(d) || d <= 0) return 0;
  if (d >= 4294967295.0) return UINT32_MAX;
  return (uint32_t)d;
}

// MSB-first bit writer with the semantics of BufferedBitWriter
// (io/bit_write.rs:45-64, 239-286): the low `k` bits of a value are emitted
// most significant first; the last byte is zero padded.  Bits accumulate in
// a 64-bit register and leave as big-endian 8-byte words.
struct MsbWriter {
  std::vector<uint8_t> out;
  uint64_t acc = 0;  // pending bits, right-aligned
  unsigned n = 0;    // number of pending bits (< 64)
  MsbWriter() { out.reserve(1 << 16); }
  inline void flush64(uint64_t w) {
    const size_t o = out.size();
    out.resize(o + 8);
    const uint64_t be = __builtin_bswap64(w);
    std::memcpy(out.data() + o, &be, 8);
  }
  inline void put(uint64_t v, unsigned k) {
    if (k == 0) return;
    if (k < 64) v &= (1ull << k) - 1;
    if (n + k < 64) {
      acc = (acc << k) | v;
      n += k;
      return;
    }
    // fill the word: take the top (64 - n) bits of v
    const unsigned take = 64 - n;
    const unsigned rest = k - take;
    const uint64_t word = (n ? (acc << take) : 0) | (take == 64 ? v : (v >> rest));
    flush64(word);
    n = rest;
    acc = rest ? (v & ((1ull << rest) - 1)) : 0;
  }
  inline void bit(bool b) { put(b ? 1u : 0u, 1); }
  std::vector<uint8_t> finish() {
    // remaining n bits, MSB first, zero padded to a byte
    while (n >= 8) {
      out.push_back((uint8_t)(acc >> (n - 8)));
      n -= 8;
    }
    if (n) out.push_back((uint8_t)(acc << (8 - n)));
    n = 0;
    acc = 0;
    return std::move(out);
  }
};

// Tables shared by Chimp, Chimp128 and ElfXor (compressor/chimp.rs:56-66).
const uint8_t kLeadRepr[64] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 4, 4, 5, 5,
                               6, 6, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7,
                               7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7, 7};
const uint8_t kLeadRound[64] = {0,  0,  0,  0,  0,  0,  0,  0,  8,  8,  8,  8,  12, 12, 12, 12,
                                16, 16, 18, 18, 20, 20, 22, 22, 24, 24, 24, 24, 24, 24, 24, 24,
                                24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24,
                                24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24, 24};

// ---------------------------------------------------------------- Gorilla
// compressor/gorilla.rs:111-190 (modified tsz): 5-bit leading zeros clamped
// at 31, 6-bit (sig - 1), sentinels L = T = 64, close = NaN then a 0 bit.
std::vector<uint8_t> enc_gorilla(const double *v, size_t n) {
  MsbWriter w;
  uint64_t prev = 0;
  unsigned L = 64, T = 64;
  auto add = [&](uint64_t b, bool first) {
    if (first) {
      w.put(b, 64);
      prev = b;
      return;
    }
    uint64_t x = b ^ prev;
    prev = b;
    if (!x) {
      w.bit(false);
      return;
    }
    w.bit(true);
    unsigned lz = clz64(x);
    if (lz >= 32) lz = 31;
    unsigned tz = ctz64(x);
    if (lz >= L && tz >= T) {
      w.bit(false);
      w.put(x >> T, 64 - L - T);
    } else {
      w.bit(true);
      w.put(lz, 5);
      unsigned sig = 64 - lz - tz;
      w.put(sig - 1, 6);
      w.put(x >> tz, sig);
      L = lz;
      T = tz;
    }
  };
  for (size_t i = 0; i < n; ++i) add(f64_bits(v[i]), i == 0);
  add(kNanBits, n == 0);
  w.bit(false);
  return w.finish();
}

// ------------------------------------------------------------------ Chimp
// compressor/chimp.rs:99-146: the first value as 64 raw bits, then a 2-bit
// flag per value (00 same, 01 lead code + sig + payload, 10 payload with the
// stored lead, 11 lead code + payload).
struct ChimpEnc {
  uint64_t stored = 0;
  uint32_t stored_lead = UINT32_MAX;
  bool first = true;
  void add(MsbWriter &w, uint64_t b) {
    if (first) {
      first = false;
      stored = b;
      w.put(b, 64);
      return;
    }
    uint64_t x = stored ^ b;
    if (!x) {
      w.put(0, 2);
      stored_lead = 65;
    } else {
      uint32_t lead = kLeadRound[clz64(x)];
      uint32_t tz = ctz64(x);
      if (tz > 6) {
        uint32_t sig = 64 - lead - tz;
        w.put(1, 2);
        w.put(kLeadRepr[lead], 3);
        w.put(sig, 6);
        w.put(x >> tz, sig);
        stored_lead = 65;
      } else if (lead == stored_lead) {
        w.put(2, 2);
        w.put(x, 64 - lead);
      } else {
        stored_lead = lead;
        w.put(3, 2);
        w.put(kLeadRepr[lead], 3);
        w.put(x, 64 - lead);
      }
    }
    stored = b;
  }
};

std::vector<uint8_t> enc_chimp(const double *v, size_t n) {
  MsbWriter w;
  ChimpEnc c;
  for (size_t i = 0; i < n; ++i) c.add(w, f64_bits(v[i]));
  c.add(w, kNanBits);  // close: the NaN end marker, then a 0 bit
  w.bit(false);
  return w.finish();
}

// --------------------------------------------------------------- Chimp128
// compressor/chimp_n.rs:196-258: 128-value ring, 14-bit key index table,
// THRESHOLD = 13, flag-00 field 9 bits, flag-01 field 18 bits.
std::vector<uint8_t> enc_chimp128(const double *v, size_t n) {
  constexpr size_t N = 128;
  constexpr uint64_t SET_LSB = (1u << 14) - 1;
  MsbWriter w;
  std::vector<uint64_t> stored(N, 0);
  std::vector<size_t> indices(size_t(1) << 14, 0);
  uint32_t stored_lead = UINT32_MAX;
  size_t index = 0, current = 0;
  auto add = [&](uint64_t b, bool first) {
    if (first) {
      stored[current] = b;
      w.put(b, 64);
      indices[b & SET_LSB] = index;
      return;
    }
    size_t key = b & SET_LSB;
    uint64_t x;
    size_t prev_idx;
    uint32_t tz = 0;
    size_t cur_idx = indices[key];
    if (index - cur_idx < N) {
      uint64_t tx = b ^ stored[cur_idx % N];
      tz = ctz64(tx);
      if (tz > 13) {
        prev_idx = cur_idx % N;
        x = tx;
      } else {
        prev_idx = index % N;
        x = stored[prev_idx] ^ b;
      }
    } else {
      prev_idx = index % N;
      x = stored[prev_idx] ^ b;
    }
    if (!x) {
      w.put(prev_idx, 9);
      stored_lead = 65;
    } else {
      uint32_t lead = kLeadRound[clz64(x)];
      if (tz > 13) {
        uint32_t sig = 64 - lead - tz;
        w.put(512ull * (N + prev_idx) + 64ull * kLeadRepr[lead] + sig, 18);
        w.put(x >> tz, sig);
        stored_lead = 65;
      } else if (lead == stored_lead) {
        w.put(2, 2);
        w.put(x, 64 - lead);
      } else {
        stored_lead = lead;
        w.put(24 + kLeadRepr[lead], 5);
        w.put(x, 64 - lead);
      }
    }
    current = (current + 1) % N;
    stored[current] = b;
    index += 1;
    indices[key] = index;
  };
  for (size_t i = 0; i < n; ++i) add(f64_bits(v[i]), i == 0);
  add(kNanBits, n == 0);
  w.bit(false);
  return w.finish();
}

// -------------------------------------------------------------------- Elf
// compression_common/elf.rs (decimal significand analysis).
const uint32_t kFAlpha[21] = {0,  4,  7,  10, 14, 17, 20, 24, 27, 30, 34,
                              37, 40, 44, 47, 50, 54, 57, 60, 64, 67};
const double kSpLess1[11] = {1.0,    0.1,     0.01,     0.001,     0.0001,     0.00001,
                             0.000001, 0.0000001, 0.00000001, 0.000000001, 0.0000000001};

// Powers of ten as the reference obtains them: exact literals below 1e21,
// otherwise the correctly rounded result of parsing "1.0E{i}"
// (compression_common/elf.rs:95-109).  glibc strtod rounds correctly.
struct Pow10 {
  double pos[400];
  double neg[400];
  Pow10() {
    char buf[32];
    for (int i = 0; i < 400; ++i) {
      std::snprintf(buf, sizeof buf, "1.0E%d", i);
      pos[i] = std::strtod(buf, nullptr);
      std::snprintf(buf, sizeof buf, "1.0E-%d", i);
      neg[i] = std::strtod(buf, nullptr);
    }
  }
};
const Pow10 &pow10_table() {
  static const Pow10 t;
  return t;
}
inline double get_10i_p(uint32_t i) { return i < 400 ? pow10_table().pos[i] : INFINITY; }

uint32_t get_f_alpha(uint32_t alpha) {
  if (alpha >= 21) return sat_u32(std::ceil((double)alpha * 3.321928094887362));
  return kFAlpha[alpha];
}

void get_sp_and_flag(double v, int32_t *sp, int32_t *flag) {
  static const uint64_t greater1[10] = {1,      10,      100,      1000,      10000,
                                        100000, 1000000, 10000000, 100000000, 1000000000};
  if (v >= 1.0) {
    for (int i = 0; i < 9; ++i)
      if (v < (double)greater1[i + 1]) {
        *sp = i;
        *flag = 0;
        return;
      }
  } else {
    for (int i = 1; i < 11; ++i)
      if (v >= kSpLess1[i]) {
        *sp = -i;
        *flag = v == kSpLess1[i] ? 1 : 0;
        return;
      }
  }
  double l = std::log10(v);
  *sp = (int32_t)std::floor(l);
  *flag = l == std::floor(l) ? 1 : 0;
}

int32_t get_significant_count(double v, int32_t sp, int32_t last_beta_star) {
  int32_t i;
  if (last_beta_star != INT32_MAX && last_beta_star != 0)
    i = std::max(last_beta_star - sp - 1, 1);
  else if (last_beta_star == INT32_MAX)
    i = 17 - sp - 1;
  else if (sp >= 0)
    i = 1;
  else
    i = -sp;
  double temp = v * get_10i_p((uint32_t)i);
  int64_t ti = sat_i64(temp);
  while ((double)ti != temp) {
    i += 1;
    temp = v * get_10i_p((uint32_t)i);
    if (std::isinf(temp)) return 17;
    ti = sat_i64(temp);
  }
  if (temp / get_10i_p((uint32_t)i) != v) return 17;
  while (i > 0 && ti % 10 == 0) {
    i -= 1;
    ti /= 10;
  }
  return sp + i + 1;
}

// ElfXor stream writer, compressor/elf.rs:186-277.
struct ElfXorEnc {
  uint64_t stored = 0;
  uint32_t L = UINT32_MAX, T = UINT32_MAX;
  bool first = true;
  void add(MsbWriter &w, uint64_t b) {
    if (first) {
      first = false;
      stored = b;
      uint32_t tz = ctz64(b);
      w.put(tz, 7);
      if (tz < 63) w.put(b >> (tz + 1), 63 - tz);
      return;
    }
    uint64_t x = stored ^ b;
    if (!x) {
      w.put(1, 2);
      return;
    }
    uint32_t lead = kLeadRound[clz64(x)];
    uint32_t tz = ctz64(x);
    if (lead == L && tz >= T) {
      uint32_t center = 64 - L - T;
      uint32_t len = 2 + center;
      if (len > 64) {
        w.put(0, 2);
        w.put(x >> T, center);
      } else {
        w.put(x >> T, len);
      }
    } else {
      L = lead;
      T = tz;
      uint32_t center = 64 - L - T;
      uint32_t repr = kLeadRepr[L];
      if (center <= 16)
        w.put((((0x2u << 3) | repr) << 4) | (center & 0xf), 9);
      else
        w.put((((0x3u << 3) | repr) << 6) | (center & 0x3f), 11);
      if (center > 1) w.put(x >> (T + 1), center - 1);
    }
    stored = b;
  }
};

// Elf wrapper, compressor/elf.rs:91-160: erase low mantissa bits when the
// decimal significand allows it and tag the value with a beta* prefix; the
// erased values go to an XOR encoder, ElfXor (Elf) or Chimp (ElfOnChimp,
// compressor/elf.rs:36), which is closed after a case-10 prefix.
template <class Xor>
std::vector<uint8_t> enc_elf_over(const double *v, size_t n) {
  MsbWriter w;
  Xor x;
  int32_t last_beta_star = INT32_MAX;
  for (size_t k = 0; k < n; ++k) {
    double d = v[k];
    uint64_t bits = f64_bits(d);
    uint64_t vp;
    if (d == 0.0 || std::isinf(d)) {
      w.put(2, 2);
      vp = bits;
    } else if (std::isnan(d)) {
      w.put(2, 2);
      vp = kNanBits;
    } else {
      double a = std::fabs(d);
      int32_t sp, flag;
      get_sp_and_flag(a, &sp, &flag);
      int32_t beta = get_significant_count(a, sp, last_beta_star);
      uint32_t alpha = (uint32_t)(beta - sp - 1);
      int32_t beta_star = flag == 1 ? 0 : beta;
      uint64_t mask = ~0ull, delta = 0;
      int32_t erase_bits = 0;
      if (beta_star < 16) {
        uint32_t e = (uint32_t)((bits >> 52) & 0x7ff);
        uint32_t g_alpha = get_f_alpha(alpha) + e - 1023u;
        erase_bits = 52 - (int32_t)g_alpha;
        if (erase_bits > 4) {
          mask = ~0ull << (erase_bits & 63);
          delta = ~mask & bits;
        }
      }
      if (beta_star < 16 && erase_bits > 4 && delta != 0) {
        if (beta_star == last_beta_star) {
          w.bit(false);
        } else {
          w.put((uint32_t)beta_star | 0x30u, 6);
          last_beta_star = beta_star;
        }
        vp = mask & bits;
      } else {
        w.put(2, 2);
        vp = bits;
      }
    }
    x.add(w, vp);
  }
  // close: case 10 prefix, then the XOR encoder's terminator (NaN and a 0 bit).
  w.put(2, 2);
  x.add(w, kNanBits);
  w.bit(false);
  return w.finish();
}
std::vector<uint8_t> enc_elf(const double *v, size_t n) { return enc_elf_over<ElfXorEnc>(v, n); }
std::vector<uint8_t> enc_elf_on_chimp(const double *v, size_t n) { return enc_elf_over<ChimpEnc>(v, n); }

// ------------------------------------------------------------------- BUFF
const uint64_t kPrecisionMap[13] = {0, 5, 8, 11, 15, 18, 21, 25, 28, 31, 35, 38, 50};

int buff_frac_bits_impl(uint64_t scale, uint32_t *frac) {
  size_t precision = 0;
  if (scale != 0) {
    float l = std::log10((float)scale);
    precision = (std::isnan(l) || l <= 0.0f) ? 0 : (size_t)l;
  }
  if (precision >= 13) return -1;
  *frac = (uint32_t)kPrecisionMap[precision];
  return 0;
}

// compressor/buff.rs:222-305: 20-byte header (base lo/hi, n, B, frac), then
// B in 1..7: n x B bits LSB-first; B >= 8: floor(B/8) byte planes (MSB plane
// first, each byte flipped by 0x80) and an LSB-first tail of B % 8 bits.
std::vector<uint8_t> enc_buff(const double *v, size_t n, uint32_t frac) {
  std::vector<int64_t> fx(n);
  int64_t mn = INT64_MAX, mx = INT64_MIN;
  for (size_t i = 0; i < n; ++i) {
    fx[i] = buff_fixed(v[i], (int32_t)frac);
    if (fx[i] < mn) mn = fx[i];
    if (fx[i] > mx) mx = fx[i];
  }
  uint64_t span = (uint64_t)mx - (uint64_t)mn;
  uint32_t B = span == 0 ? 0 : 64 - clz64(span);
  std::vector<uint8_t> out;
  put_u32(out, (uint32_t)(uint64_t)mn);
  put_u32(out, (uint32_t)((uint64_t)mn >> 32));
  put_u32(out, (uint32_t)n);
  put_u32(out, B);
  put_u32(out, frac);
  auto pack_lsb = [&](unsigned r) {
    size_t start = out.size();
    out.resize(start + (n * r + 7) / 8, 0);
    uint8_t *p = out.data() + start;
    uint64_t bitpos = 0;
    for (size_t i = 0; i < n; ++i) {
      uint32_t t = (uint32_t)(((uint64_t)fx[i] - (uint64_t)mn) & ((1u << r) - 1));
      for (unsigned k = 0; k < r; ++k, ++bitpos)
        if (t >> k & 1) p[bitpos >> 3] |= (uint8_t)(1u << (bitpos & 7));
    }
  };
  if (B == 0) return out;
  if (B < 8) {
    pack_lsb(B);
    return out;
  }
  unsigned rem = B;
  while (rem >= 8) {
    rem -= 8;
    size_t start = out.size();
    out.resize(start + n);
    for (size_t i = 0; i < n; ++i)
      out[start + i] = (uint8_t)((((uint64_t)fx[i] - (uint64_t)mn) >> rem) ^ 0x80u);
  }
  if (rem > 0) pack_lsb(rem);
  return out;
}

// ------------------------------------------------------------ DeltaSprintz
// compressor/sprintz.rs:151-205: q = round(v * scale) as i64, zigzag deltas
// (first delta 0), header q0 (64 bits) | scale (32 bits) | W (8 bits), then
// n x W bits, all MSB-first.
inline uint64_t zigzag(int64_t d) { return (uint64_t)((d << 1) ^ (d >> 63)); }

std::vector<uint8_t> enc_sprintz(const double *v, size_t n, uint64_t scale) {
  std::vector<uint64_t> zz(n);
  int64_t q0 = n ? sat_i64(std::round(v[0] * (double)scale)) : 0;
  int64_t prev = q0;
  uint64_t acc = 0;
  for (size_t i = 0; i < n; ++i) {
    int64_t q = sat_i64(std::round(v[i] * (double)scale));
    zz[i] = zigzag((int64_t)((uint64_t)q - (uint64_t)prev));
    acc |= zz[i];
    prev = q;
  }
  unsigned W = 64 - clz64(acc);
  MsbWriter w;
  w.put((uint64_t)q0, 64);
  w.put(scale, 32);
  w.put(W, 8);
  if (W)
    for (size_t i = 0; i < n; ++i) w.put(zz[i], W);
  return w.finish();
}

// -------------------------------------------------------------------- RLE
// compressor/run_length.rs:64-92: u64 field count, then (u8 run, f64) fields;
// runs compare bit patterns and split at 255.
std::vector<uint8_t> enc_rle(const double *v, size_t n) {
  std::vector<uint8_t> out(8, 0);
  uint64_t fields = 0;
  auto add = [&](uint64_t bits, uint8_t run) {
    out.push_back(run);
    put_u64(out, bits);
    ++fields;
  };
  if (n) {
    uint64_t pv = f64_bits(v[0]);
    unsigned run = 1;
    for (size_t i = 1; i < n; ++i) {
      uint64_t b = f64_bits(v[i]);
      if (b == pv && run < 255) {
        ++run;
        continue;
      }
      add(pv, (uint8_t)run);
      pv = b;
      run = 1;
    }
    add(pv, (uint8_t)run);
  }
  std::memcpy(out.data(), &fields, 8);
  return out;
}

std::vector<uint8_t> enc_uncompressed(const double *v, size_t n) {
  std::vector<uint8_t> out(n * 8);
  if (n) std::memcpy(out.data(), v, n * 8);
  return out;
}

int encode_chunk(int scheme, const double *v, size_t n, const ebi_encoder_config &cfg,
                 uint32_t buff_frac, std::vector<uint8_t> *out) {
  switch (scheme) {
    case kUncompressed: *out = enc_uncompressed(v, n); return 0;
    case kRLE: *out = enc_rle(v, n); return 0;
    case kGorilla: *out = enc_gorilla(v, n); return 0;
    case kChimp: *out = enc_chimp(v, n); return 0;
    case kChimp128: *out = enc_chimp128(v, n); return 0;
    case kElf: *out = enc_elf(v, n); return 0;
    case kElfOnChimp: *out = enc_elf_on_chimp(v, n); return 0;
    case kBUFF: *out = enc_buff(v, n, buff_frac); return 0;
    case kDeltaSprintz: *out = enc_sprintz(v, n, cfg.scale); return 0;
    default: return EBI_UNSUPPORTED;
  }
}

}  // namespace

int buff_frac_bits(uint64_t scale, uint32_t *frac) { return buff_frac_bits_impl(scale, frac); }

int encode_geometry(const ebi_encoder_config &cfg, size_t n, EncodeGeom *g) {
  // Chunk partition, encoder.rs:317-346.
  size_t per_chunk;
  switch (cfg.chunk_kind) {
    case EBI_CHUNK_FULL: per_chunk = n ? n : 1; break;
    case EBI_CHUNK_RECORD_COUNT: per_chunk = (size_t)cfg.chunk_value; break;
    case EBI_CHUNK_BYTE_SIZE:
      per_chunk = (size_t)std::ceil(((double)cfg.chunk_value / 8.0) * 1.5);
      break;
    default: return EBI_INVALID_ARGUMENT;
  }
  if (per_chunk == 0) return EBI_INVALID_ARGUMENT;
  g->per_chunk = per_chunk;
  g->n_chunks = (n + per_chunk - 1) / per_chunk;
  return EBI_OK;
}

size_t encoded_file_size(int scheme, uint64_t body, size_t n_chunks) {
  return kFileHeaderSize + body + kFooter0Size + n_chunks * kChunkFooterSize + 1 + config_size_of(scheme) +
         kFooter3Size;
}

void write_header_footer(uint8_t *buf, const ebi_encoder_config &cfg, size_t n, const EncodeGeom &g,
                         const std::vector<uint64_t> &phys) {
  const int scheme = cfg.scheme;
  const uint64_t body = phys[g.n_chunks] - kFileHeaderSize;
  const int cfg_size = config_size_of(scheme);
  // FileHeader, written with footer_offset already known (the reference
  // back-patches the same slot, encoder.rs:236-241).
  std::vector<uint8_t> h;
  h.insert(h.end(), {'E', 'B', 'I', '1'});
  h.push_back(0);
  h.push_back((uint8_t)scheme);
  h.push_back((uint8_t)cfg.chunk_kind);
  h.push_back(0);
  put_u64(h, cfg.chunk_kind == EBI_CHUNK_FULL ? 0 : cfg.chunk_value);
  put_u64(h, kFileHeaderSize + body);
  for (int i = 0; i < 3; ++i) put_u16(h, kVersion[i]);
  std::memcpy(buf, h.data(), h.size());
  std::vector<uint8_t> f;
  f.reserve(kFooter0Size + g.n_chunks * kChunkFooterSize + 1 + cfg_size + kFooter3Size);
  put_u64(f, n);
  put_u64(f, g.n_chunks);
  for (size_t c = 0; c < g.n_chunks; ++c) {
    put_u64(f, phys[c]);
    put_u64(f, c * g.per_chunk);
  }
  f.push_back((uint8_t)cfg_size);
  const uint64_t capacity = cfg.capacity ? cfg.capacity : kDefaultCapacity;
  if (scheme == kBUFF) {
    put_u64(f, cfg.scale);
  } else if (scheme == kDeltaSprintz) {
    put_u64(f, capacity);
    put_u64(f, cfg.scale);
  } else {
    put_u64(f, capacity);
  }
  // FileFooter3: compression time, 9 phase timers, crc -- all zero here so
  // that files are byte-reproducible (the reference stores wall-clock times).
  f.resize(f.size() + kFooter3Size, 0);
  std::memcpy(buf + kFileHeaderSize + body, f.data(), f.size());
}

// ----------------------------------------------------------------- parser
int parse_file(const uint8_t *file, size_t len, FileMeta *m, std::string *err) {
  if (len < kFileHeaderSize) {
    *err = "file shorter than the 30-byte header";
    return EBI_UNEXPECTED_END;
  }
  if (file[4] != 0) {
    *err = "unknown field type";
    return EBI_CONVERSION;
  }
  if (file[5] >= kSchemeCount) {
    *err = "unknown compression scheme";
    return EBI_CONVERSION;
  }
  m->scheme = file[5];
  m->chunk_kind = file[6];
  m->chunk_value = rd_u64(file + 8);
  m->footer_offset = rd_u64(file + 16);
  for (int i = 0; i < 3; ++i) m->version[i] = rd_u16(file + 24 + 2 * i);
  uint64_t fo = m->footer_offset;
  if (fo < kFileHeaderSize || fo > len || len - fo < kFooter0Size) {
    *err = "footer offset outside the file";
    return EBI_UNEXPECTED_END;
  }
  m->number_of_records = rd_u64(file + fo);
  uint64_t n_chunks = rd_u64(file + fo + 8);
  uint64_t pos = fo + kFooter0Size;
  if (n_chunks > (len - pos) / kChunkFooterSize) {
    *err = "chunk footers run past the end of the file";
    return EBI_UNEXPECTED_END;
  }
  std::vector<uint64_t> phys(n_chunks), logi(n_chunks);
  for (uint64_t c = 0; c < n_chunks; ++c) {
    phys[c] = rd_u64(file + pos);
    logi[c] = rd_u64(file + pos + 8);
    pos += kChunkFooterSize;
  }
  if (pos + 1 > len) {
    *err = "missing config size byte";
    return EBI_UNEXPECTED_END;
  }
  size_t cfg_size = file[pos++];
  if (pos + cfg_size + kFooter3Size > len) {
    *err = "config / footer3 run past the end of the file";
    return EBI_UNEXPECTED_END;
  }
  m->capacity = 0;
  m->scale = 0;
  if (m->scheme == kBUFF && cfg_size >= 8) m->scale = rd_u64(file + pos);
  if (m->scheme == kDeltaSprintz && cfg_size >= 16) {
    m->capacity = rd_u64(file + pos);
    m->scale = rd_u64(file + pos + 8);
  } else if (m->scheme != kBUFF && cfg_size >= 8) {
    m->capacity = rd_u64(file + pos);
  }
  m->chunks.resize(n_chunks);
  for (uint64_t c = 0; c < n_chunks; ++c) {
    uint64_t next_phys = c + 1 < n_chunks ? phys[c + 1] : fo;
    uint64_t next_log = c + 1 < n_chunks ? logi[c + 1] : m->number_of_records;
    if (phys[c] < kFileHeaderSize || next_phys < phys[c] || next_phys > fo) {
      *err = "chunk byte range outside the chunk area";
      return EBI_UNEXPECTED_END;
    }
    if (next_log < logi[c]) {
      *err = "chunk logical offsets decrease";
      return EBI_CONVERSION;
    }
    m->chunks[c] = {phys[c], next_phys - phys[c], logi[c], next_log - logi[c]};
  }
  return EBI_OK;
}

}  // namespace ebi

using namespace ebi;

extern "C" int ebi_encode(const double *values, size_t n, const ebi_encoder_config *cfg,
                          uint8_t **out, size_t *out_len) {
  if (!cfg || !out || !out_len || (!values && n)) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  *out_len = 0;
  int scheme = cfg->scheme;
  if (scheme < 0 || scheme > kDeltaSprintz) return EBI_UNSUPPORTED;
  uint32_t frac = 0;
  if (scheme == kBUFF && buff_frac_bits(cfg->scale, &frac) != 0) return EBI_INVALID_ARGUMENT;

  EncodeGeom geom;
  if (int rc = encode_geometry(*cfg, n, &geom)) return rc;
  const size_t per_chunk = geom.per_chunk;
  const size_t n_chunks = geom.n_chunks;
  std::vector<std::vector<uint8_t>> bodies(n_chunks);
  std::atomic<size_t> next{0};
  std::atomic<int> failed{0};
  auto work = [&]() {
    for (;;) {
      size_t c = next.fetch_add(1);
      if (c >= n_chunks) return;
      size_t lo = c * per_chunk, hi = std::min(n, lo + per_chunk);
      int rc = encode_chunk(scheme, values + lo, hi - lo, *cfg, frac, &bodies[c]);
      if (rc) failed = rc;
    }
  };
  int nt = std::max(1, std::min<int>(cfg->n_threads, 64));
  nt = (int)std::min<size_t>((size_t)nt, std::max<size_t>(n_chunks, 1));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  if (failed) return failed;

  std::vector<uint64_t> phys(n_chunks + 1);
  phys[0] = kFileHeaderSize;
  for (size_t c = 0; c < n_chunks; ++c) phys[c + 1] = phys[c] + bodies[c].size();
  const uint64_t body = phys[n_chunks] - kFileHeaderSize;
  const size_t total = encoded_file_size(scheme, body, n_chunks);
  uint8_t *buf = (uint8_t *)std::malloc(total);
  if (!buf) return EBI_OOM;
  // chunk bodies, copied in parallel
  next = 0;
  auto copy = [&]() {
    for (;;) {
      size_t c = next.fetch_add(1);
      if (c >= n_chunks) return;
      if (!bodies[c].empty()) std::memcpy(buf + phys[c], bodies[c].data(), bodies[c].size());
      std::vector<uint8_t>().swap(bodies[c]);
    }
  };
  pool.clear();
  for (int t = 1; t < nt; ++t) pool.emplace_back(copy);
  copy();
  for (auto &t : pool) t.join();
  write_header_footer(buf, *cfg, n, geom, phys);
  *out = buf;
  *out_len = total;
  return EBI_OK;
}

extern "C" void ebi_buffer_free(void *p) { std::free(p); }

extern "C" int ebi_file_parse(const uint8_t *file, size_t len, ebi_file_info *info) {
  if (!file || !info) return EBI_INVALID_ARGUMENT;
  FileMeta m;
  std::string err;
  int rc = parse_file(file, len, &m, &err);
  if (rc) return rc;
  info->number_of_records = m.number_of_records;
  info->number_of_chunks = m.chunks.size();
  info->footer_offset = m.footer_offset;
  info->chunk_option_value = m.chunk_value;
  info->config_capacity = m.capacity;
  info->config_scale = m.scale;
  info->scheme = m.scheme;
  info->chunk_option_kind = m.chunk_kind;
  for (int i = 0; i < 3; ++i) info->version[i] = m.version[i];
  info->reserved = 0;
  return EBI_OK;
}

extern "C" int ebi_file_chunks(const uint8_t *file, size_t len, uint64_t *physical_offset,
                               uint64_t *chunk_size, uint64_t *logical_offset,
                               uint64_t *number_of_records) {
  if (!file) return EBI_INVALID_ARGUMENT;
  FileMeta m;
  std::string err;
  int rc = parse_file(file, len, &m, &err);
  if (rc) return rc;
  for (size_t c = 0; c < m.chunks.size(); ++c) {
    if (physical_offset) physical_offset[c] = m.chunks[c].physical_offset;
    if (chunk_size) chunk_size[c] = m.chunks[c].size;
    if (logical_offset) logical_offset[c] = m.chunks[c].logical_offset;
    if (number_of_records) number_of_records[c] = m.chunks[c].number_of_records;
  }
  return EBI_OK;
}
