// Host check of the Elf helpers: elf_get_sp (threshold counts) against the
// reference's table walk of compression_common/elf.rs:111-130 over tables
// parsed with strtod exactly as the reference builds them (elf.rs:95-109),
// on every power of ten and its two neighbours, special values and random
// magnitudes; p10_exact(a) against the parsed "1.0E{a}" for a <= 22; the
// table form of the Elf value recovery (elf_recover_tab) against
// elf_recover, and its division (div_rn) against IEEE division, for the
// powers of ten and for DeltaSprintz scales.
// Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <vector>

#include "../../ebi_amd/csrc/kernels/codecs.hpp"

static double p10[400], n10[400];

static int32_t walk_sp(double v) {  // the reference's loop, over the parsed tables
  if (v >= 1.0) {
    for (int i = 0; i < 9; ++i)
      if (v < p10[i + 1]) return i;
  } else {
    for (int i = 1; i < 11; ++i)
      if (v >= n10[i]) return -i;
  }
  return ebi::sat_i32(std::floor(std::log10(v)));
}

int main() {
  char buf[32];
  for (int i = 0; i < 400; ++i) {
    std::snprintf(buf, sizeof buf, "1.0E%d", i);
    p10[i] = std::strtod(buf, nullptr);
    std::snprintf(buf, sizeof buf, "1.0E-%d", i);
    n10[i] = std::strtod(buf, nullptr);
  }
  long fails = 0, checks = 0;
  for (uint32_t a = 0; a <= 22; ++a, ++checks)
    if (ebi::p10_exact(a) != p10[a]) {
      std::printf("p10_exact(%u) = %.17g, parsed %.17g\n", a, ebi::p10_exact(a), p10[a]);
      ++fails;
    }
  std::vector<double> vs = {0.0, 1.0, std::numeric_limits<double>::denorm_min(), std::numeric_limits<double>::min(),
                            std::numeric_limits<double>::max(), std::numeric_limits<double>::infinity(),
                            std::numeric_limits<double>::quiet_NaN()};
  for (int i = 0; i < 40; ++i)
    for (double x : {p10[i], n10[i]}) {
      vs.push_back(x);
      vs.push_back(std::nextafter(x, 0.0));
      vs.push_back(std::nextafter(x, std::numeric_limits<double>::infinity()));
    }
  std::mt19937_64 rng(2024);
  std::uniform_real_distribution<double> e(-14.0, 14.0);
  for (int i = 0; i < 1000000; ++i) vs.push_back(std::pow(10.0, e(rng)));
  for (double v : vs) {
    ++checks;
    const int32_t a = ebi::elf_get_sp(v), b = walk_sp(v);
    if (a != b) {
      if (fails < 10) std::printf("get_sp(%.17g) = %d, table walk %d\n", v, a, b);
      ++fails;
    }
  }
  // the table form of the value recovery against elf_recover, on the same
  // magnitudes with both signs and every beta*, over the library's tables
  for (int i = 0; i < 400; ++i) {
    ebi::c_p10[i] = p10[i];
    ebi::c_n10[i] = n10[i];
  }
  double tab[ebi::kElfTab];
  for (uint32_t i = 0; i < ebi::kElfTab; ++i) ebi::elf_tab_fill(tab, i, p10, n10);
  for (double x : vs)
    for (double v : {x, -x})
      for (int32_t beta = 0; beta < 16; ++beta) {
        ++checks;
        const double a = ebi::elf_recover(v, beta), b = ebi::elf_recover_tab(v, beta, tab);
        if (ebi::as_u64(a) != ebi::as_u64(b) && !(a != a && b != b)) {
          if (fails < 10) std::printf("recover(%.17g, %d) = %.17g, table form %.17g\n", v, beta, a, b);
          ++fails;
        }
      }
  // div_rn against IEEE division: every integer up to 2e5, random
  // integers below 2^53, random integer-valued doubles up to 1e31
  std::uniform_int_distribution<uint64_t> i53(1, (1ull << 53) - 1);
  std::uniform_real_distribution<double> big(0.0, 31.0 * 3.321928094887362);
  for (uint32_t a = 0; a <= 22; ++a) {
    const double D = p10[a], y = n10[a];
    auto one = [&](double N) {
      for (double M : {N, -N}) {
        ++checks;
        const double q = ebi::div_rn(M, D, y), w = M / D;
        if (ebi::as_u64(q) != ebi::as_u64(w)) {
          if (fails < 10) std::printf("div_rn(%.17g, 1e%u) = %.17g, IEEE %.17g\n", M, a, q, w);
          ++fails;
        }
      }
    };
    for (uint32_t N = 1; N <= 200000; ++N) one((double)N);
    for (int i = 0; i < 200000; ++i) one((double)i53(rng));
    for (int i = 0; i < 200000; ++i) {
      const double N = std::floor(std::exp2(big(rng)));
      if (N >= 1.0) one(N);
    }
  }
  // div_rn for the DeltaSprintz scale: random u32 scales (and the small and
  // power-of-ten ones) against IEEE division of random integers of every
  // magnitude up to 2^63, both signs
  std::uniform_int_distribution<uint32_t> u32(1, 0xffffffffu);
  std::uniform_int_distribution<int> mag(0, 63);
  std::vector<double> scales = {1, 2, 3, 7, 10, 100, 1000, 10000, 65535, 4294967295.0};
  for (int i = 0; i < 200; ++i) scales.push_back((double)u32(rng));
  for (double D : scales) {
    const double y = 1.0 / D;
    for (int i = 0; i < 20000; ++i) {
      const int m = mag(rng);
      const int64_t iN = (int64_t)(rng() >> (63 - m));
      for (double N : {(double)iN, (double)(-iN)}) {  // int64 -> f64: never -0
        ++checks;
        const double q = ebi::div_rn(N, D, y), w = N / D;
        if (ebi::as_u64(q) != ebi::as_u64(w)) {
          if (fails < 10) std::printf("div_rn(%.17g, %.17g) = %.17g, IEEE %.17g\n", N, D, q, w);
          ++fails;
        }
      }
    }
  }
  std::printf("%s: %ld checks, %ld failures\n", fails ? "FAIL" : "OK", checks, fails);
  return fails ? 1 : 0;
}
