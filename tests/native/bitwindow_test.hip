// Host check of BitWindow, the 128-bit MSB-first window the XOR decoders
// (Gorilla, Chimp, Chimp128, Elf) read their streams through: random field
// extractions, consumes and refills against a bit-by-bit reader.  Exit code
// 0 on success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../ebi_amd/csrc/kernels/common.hpp"

static uint64_t naive(const uint8_t *b, uint64_t pos, uint32_t len) {
  uint64_t v = 0;
  for (uint32_t i = 0; i < len; ++i) {
    const uint64_t p = pos + i;
    v = (v << 1) | ((b[p >> 3] >> (7 - (p & 7))) & 1u);
  }
  return v;
}

int main() {
  const size_t bytes = 1 << 16;
  uint8_t *buf = static_cast<uint8_t *>(std::aligned_alloc(256, bytes));
  std::mt19937_64 rng(12345);
  for (size_t i = 0; i < bytes; ++i) buf[i] = (uint8_t)rng();
  long fails = 0, checks = 0;
  for (int trial = 0; trial < 4000; ++trial) {
    const uint32_t start = (uint32_t)(rng() % 8192);
    ebi::BitWindow r;
    r.init(buf, start);
    uint64_t pos = start;
    for (int step = 0; step < 200; ++step) {
      if (r.nb < 96 || r.pos != pos) ++fails;
      const uint32_t o = (uint32_t)(rng() % 32), len = 1 + (uint32_t)(rng() % 64);
      ++checks;
      if (r.field(o, len) != naive(buf, pos + o, len)) ++fails;
      if (r.field(0, 64) != naive(buf, pos, 64)) ++fails;
      const uint32_t k = (uint32_t)(rng() % 97);
      r.consume(k);
      pos += k;
      r.refill();
    }
  }
  std::free(buf);
  std::printf("%s: %ld checks, %ld failures\n", fails ? "FAIL" : "OK", checks, fails);
  return fails ? 1 : 0;
}
