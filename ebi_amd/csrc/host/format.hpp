// EBI1 on-disk layout helpers shared by the host side of libebi_hip.
//
// Byte layout (all little-endian, packed), following ebi/src/format.rs:143-258:
//   FileHeader (30 B): magic "EBI1" | field_type u8 | scheme u8 |
//                      chunk_option.kind u8 | reserved u8 | chunk_option.value u64 |
//                      footer_offset u64 | version u16 x3
//   chunks, contiguous from offset 30 (GeneralChunkHeader is zero-sized)
//   footer @footer_offset: number_of_records u64 | number_of_chunks u64 |
//                          C x {physical_offset u64, logical_offset u64} |
//                          config_size u8 | config bytes |
//                          FileFooter3 (u128 + 9 x u128 + u32 crc = 164 B)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/ebi_hip.h"

namespace ebi {

constexpr size_t kFileHeaderSize = 30;
constexpr size_t kFooter0Size = 16;
constexpr size_t kChunkFooterSize = 16;
constexpr size_t kFooter3Size = 16 + 9 * 16 + 4;
constexpr uint64_t kDefaultCapacity = 8192;  // compressor.rs:465
constexpr uint16_t kVersion[3] = {0, 1, 14};  // ebi/Cargo.toml:3
constexpr uint64_t kNanBits = 0x7ff8000000000000ull;  // f64::NAN.to_bits()

enum Scheme : int {
  kUncompressed = 0,
  kRLE = 1,
  kBUFF = 2,
  kGorilla = 3,
  kChimp = 4,
  kChimp128 = 5,
  kElfOnChimp = 6,
  kElf = 7,
  kDeltaSprintz = 8,
  kSchemeCount = 13,  // ids 9..12 are zstd, gzip, snappy, ALP
};

inline uint64_t rd_u64(const uint8_t *p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}
inline uint32_t rd_u32(const uint8_t *p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return v;
}
inline uint16_t rd_u16(const uint8_t *p) {
  uint16_t v;
  std::memcpy(&v, p, 2);
  return v;
}
inline void put_u64(std::vector<uint8_t> &b, uint64_t v) {
  uint8_t t[8];
  std::memcpy(t, &v, 8);
  b.insert(b.end(), t, t + 8);
}
inline void put_u32(std::vector<uint8_t> &b, uint32_t v) {
  uint8_t t[4];
  std::memcpy(t, &v, 4);
  b.insert(b.end(), t, t + 4);
}
inline void put_u16(std::vector<uint8_t> &b, uint16_t v) {
  uint8_t t[2];
  std::memcpy(t, &v, 2);
  b.insert(b.end(), t, t + 2);
}
inline uint64_t f64_bits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return u;
}
inline double bits_f64(uint64_t u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

// Serialized config blob size (without the leading size byte) per scheme,
// compressor.rs:328-365: capacity u64 for the XOR/RLE/plain codecs, scale u64
// for BUFF, capacity+scale for DeltaSprintz, i32/u32 for zstd/gzip, 0 else.
inline int config_size_of(int scheme) {
  switch (scheme) {
    case kUncompressed: case kRLE: case kGorilla: case kChimp: case kChimp128:
    case kElfOnChimp: case kElf: case kBUFF:
      return 8;
    case kDeltaSprintz:
      return 16;
    case 9: case 10:
      return 4;
    default:
      return 0;
  }
}

// BUFF fixed-point conversion, compression_common/buff.rs:10-52: the value's
// mantissa shifted to `frac` fractional bits (truncating), negated for
// negative values; 0 below 2^-frac.  Shift amounts wrap as in the reference's
// release build.
__host__ __device__ inline int64_t buff_fixed(double v, int32_t frac) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int32_t exp = (int32_t)((b >> 52) & 0x7ff) - 1023;
  if (exp < -frac) return 0;
  const int32_t sh = 63 - exp - frac;
  uint64_t fixed = sh >= 64 ? 0 : ((b << 11) | (1ull << 63)) >> ((uint64_t)(int64_t)sh & 63);
  if (b >> 63) fixed = ~(fixed - 1);
  return (int64_t)fixed;
}

struct ChunkGeom {
  uint64_t physical_offset;
  uint64_t size;            // next physical offset (or footer_offset) - physical_offset
  uint64_t logical_offset;
  uint64_t number_of_records;
};

struct FileMeta {
  int scheme = 0;
  int chunk_kind = 0;
  uint64_t chunk_value = 0;
  uint64_t footer_offset = 0;
  uint16_t version[3] = {0, 0, 0};
  uint64_t number_of_records = 0;
  uint64_t capacity = 0;
  uint64_t scale = 0;
  std::vector<ChunkGeom> chunks;
};

// Returns 0 on success or an ebi_status code; *err receives a message.
int parse_file(const uint8_t *file, size_t len, FileMeta *meta, std::string *err);

// Write-path helpers shared by the host encoder and the GPU encoder.
struct EncodeGeom {
  size_t per_chunk = 0;  // records per chunk (the last chunk may hold fewer)
  size_t n_chunks = 0;
};
// The chunk partition of encoder.rs:317-346 (0 or an ebi_status).
int encode_geometry(const ebi_encoder_config &cfg, size_t n, EncodeGeom *g);
// Total file bytes for chunk bodies of `body` bytes.
size_t encoded_file_size(int scheme, uint64_t body, size_t n_chunks);
// Writes the FileHeader at buf and the footer after the chunk bodies, which
// start at phys[0] = 30 and end at phys[n_chunks].
void write_header_footer(uint8_t *buf, const ebi_encoder_config &cfg, size_t n, const EncodeGeom &g,
                         const std::vector<uint64_t> &phys);
// BUFF fractional bits for a scale (compressor/buff.rs:131-135), 0 or an ebi_status.
int buff_frac_bits(uint64_t scale, uint32_t *frac);

}  // namespace ebi
