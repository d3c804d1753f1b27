// Seeded synthetic columns for tests and benchmarks (not part of the query
// path).  One counter-based SplitMix64 stream per (seed, stream id) gives
// every value as a function of its index, so columns are identical for any
// thread count; normals come from Box-Muller.  Decimal rounding is
// round-half-away-from-zero, Rust's f64::round used by the reference tests
// (lemolatoon/ebi ebi/tests/aggregation.rs:14-16).
//
// Kinds (the four column generators of the plain-decode workload):
//   0 walk  -- x0 + cumsum(sigma * N(0,1)), rounded to `decimals`
//   1 unif  -- lo + (hi - lo) * U[0,1), rounded
//   2 step  -- piecewise-constant levels, run lengths ~ Geom(p), levels a
//              2-decimal random walk around 100
//   3 sin   -- 100 sin(2 pi i / period) + noise * N(0,1), rounded
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../../include/ebi_hip.h"

namespace {

inline uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t stream_seed(uint64_t seed, uint64_t stream) {
  return seed * 0x100000001B3ull + stream * 0x632BE59BD9B4E019ull;
}
// z_i = mix(s + (i + 1) * golden)
inline uint64_t sm64(uint64_t s, uint64_t i) { return mix(s + (i + 1) * 0x9E3779B97F4A7C15ull); }
inline double u01(uint64_t s, uint64_t i) { return (double)(sm64(s, i) >> 11) * (1.0 / 9007199254740992.0); }
inline double gauss(uint64_t seed, uint64_t stream, uint64_t i) {
  double u1 = u01(stream_seed(seed, 2 * stream + 1), i);
  const double u2 = u01(stream_seed(seed, 2 * stream + 2), i);
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}
inline double round_away(double x, double s) {
  const double y = x * s;
  return std::copysign(std::floor(std::fabs(y) + 0.5), y) / s + 0.0;
}

template <class F>
void parallel_for(uint64_t n, int threads, F f) {
  threads = std::max(1, std::min(threads, 64));
  if (threads == 1 || n < (1u << 16)) {
    f(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const uint64_t per = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const uint64_t a = t * per, b = std::min(n, a + per);
    if (a < b) pool.emplace_back([=] { f(a, b); });
  }
  for (auto &t : pool) t.join();
}

}  // namespace

extern "C" int ebi_synth_column(int kind, uint64_t n, uint64_t seed, int decimals, double p0,
                                double p1, double p2, int threads, double *out) {
  if (!out && n) return EBI_INVALID_ARGUMENT;
  const double s = std::pow(10.0, decimals);
  switch (kind) {
    case 0: {  // walk: p0 = x0, p1 = sigma
      parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) out[i] = p1 * gauss(seed, 0, i);
      });
      double acc = p0;
      for (uint64_t i = 0; i < n; ++i) {
        acc += out[i];
        out[i] = acc;
      }
      parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) out[i] = round_away(out[i], s);
      });
      return EBI_OK;
    }
    case 1:  // unif: p0 = lo, p1 = hi
      parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
        const uint64_t ss = stream_seed(seed, 0);
        for (uint64_t i = a; i < b; ++i) out[i] = round_away(p0 + (p1 - p0) * u01(ss, i), s);
      });
      return EBI_OK;
    case 2: {  // step: p0 = change probability
      const uint64_t ss = stream_seed(seed, 7);
      double level = 100.0;
      uint64_t changes = 0;
      for (uint64_t i = 0; i < n; ++i) {
        if (i == 0 || u01(ss, i) < p0) {
          level += gauss(seed, 9, changes++);
        }
        out[i] = level;
      }
      parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i) out[i] = round_away(out[i], s);
      });
      return EBI_OK;
    }
    case 3:  // sin: p0 = period, p1 = noise
      parallel_for(n, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; ++i)
          out[i] = round_away(100.0 * std::sin(2.0 * M_PI * (double)i / p0) + p1 * gauss(seed, 0, i), s);
      });
      return EBI_OK;
    default:
      return EBI_INVALID_ARGUMENT;
  }
}
