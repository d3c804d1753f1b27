// Host run of the XOR decoders (Gorilla, Chimp, Elf, and the Chimp128
// parse) on files from this library's encoder.  Every chunk is decoded
// whole from the stream start, and in the segment kernel's form: the index
// pass writes a checkpoint every kSeg records, then each segment restarts
// from its checkpoint and decodes in runs of kRun records with the state
// carried between runs.  Both must give back the encoded values bit for bit
// (the codecs are lossless); Chimp128's (payload, source distance) parse is
// resolved on the host the way k_chimp128_resolve defines the values.
// Synthetic columns of the four generators, ragged chunkings down to one
// record per chunk.  Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ebi_hip.h"
#include "../../ebi_amd/csrc/kernels/codecs.hpp"

using namespace ebi;

static long fails = 0, checks = 0;

static void expect(bool ok, const char *what, int codec, int gen, uint64_t chunk, uint64_t k, uint64_t i) {
  ++checks;
  if (!ok && fails++ < 20)
    std::printf("FAIL %s: codec %d gen %d chunk option %llu, chunk %llu, record %llu\n", what, codec, gen,
                (unsigned long long)chunk, (unsigned long long)k, (unsigned long long)i);
}

int main() {
  char buf[32];
  for (int i = 0; i < 400; ++i) {  // the host copies of the tables, filled as the library fills the device's
    std::snprintf(buf, sizeof buf, "1.0E%d", i);
    c_p10[i] = std::strtod(buf, nullptr);
    std::snprintf(buf, sizeof buf, "1.0E-%d", i);
    c_n10[i] = std::strtod(buf, nullptr);
  }
  struct Gen {
    int kind, dec;
    double p0, p1;
  } gens[] = {{0, 2, 100.0, 0.1}, {1, 3, -1000.0, 1000.0}, {2, 2, 0.01, 0.0}, {3, 4, 1000.0, 0.01}};
  const int codecs[] = {kCodecGorilla, kCodecChimp, kCodecElf, kCodecChimp128};
  const uint64_t chunks[] = {1, 3, 5, 64, 65, 127, 1003, 4096};
  const uint64_t n = 10003;
  std::vector<double> v(n);
  for (int codec : codecs)
    for (int g = 0; g < 4; ++g) {
      if (ebi_synth_column(gens[g].kind, n, 7 + g, gens[g].dec, gens[g].p0, gens[g].p1, 0.0, 1, v.data()) != EBI_OK)
        return 2;
      for (uint64_t chunk : chunks) {
        ebi_encoder_config cfg{};
        cfg.scheme = codec;
        cfg.chunk_kind = EBI_CHUNK_RECORD_COUNT;
        cfg.chunk_value = chunk;
        uint8_t *file = nullptr;
        size_t len = 0;
        if (ebi_encode(v.data(), n, &cfg, &file, &len) != EBI_OK) return 3;
        ebi_file_info info{};
        if (ebi_file_parse(file, len, &info) != EBI_OK) return 4;
        const size_t nc = info.number_of_chunks;
        std::vector<uint64_t> off(nc), size(nc), loff(nc), nrec(nc);
        if (ebi_file_chunks(file, len, off.data(), size.data(), loff.data(), nrec.data()) != EBI_OK) return 5;
        for (size_t k = 0; k < nc; ++k) {
          const uint32_t cn = (uint32_t)nrec[k], bits = (uint32_t)(size[k] * 8);
          // the chunk body as the device holds it: 16-byte aligned, zero padding after
          std::vector<uint4> body((size[k] + 256 + 15) / 16, make_uint4(0, 0, 0, 0));
          std::memcpy(body.data(), file + off[k], size[k]);
          const uint8_t *data = reinterpret_cast<const uint8_t *>(body.data());
          const double *want = v.data() + loff[k];
          std::vector<double> out(cn + 1, 0.0);
          const DirectSink sink{out.data()};
          const uint32_t nseg = codec == kCodecChimp128 ? cn / kSeg + 1 : (cn + kSeg - 1) / kSeg;
          std::vector<XorCkpt> ck(nseg + 1);
          XorState s;
          if (codec != kCodecChimp128) {
            xor_state_start(s, codec, data);
            uint32_t e = xor_run<true>(codec, s, bits, cn, 0, cn, sink, nullptr);
            expect(e == kErrNone, "whole-chunk decode error", codec, g, chunk, k, 0);
            for (uint32_t i = 0; i < cn; ++i)
              expect(as_u64(out[i]) == as_u64(want[i]), "whole-chunk value", codec, g, chunk, k, i);
            xor_state_start(s, codec, data);
            e = xor_run<false>(codec, s, bits, cn, 0, cn, sink, ck.data());
            expect(e == kErrNone, "index pass error", codec, g, chunk, k, 0);
            std::fill(out.begin(), out.end(), 0.0);
            for (uint32_t seg = 0; seg < nseg; ++seg) {
              const uint32_t i0 = seg * kSeg, i1 = std::min(i0 + kSeg, cn);
              xor_state_at(s, codec, data, ck[seg]);
              for (uint32_t k0 = 0; k0 < kSeg && e == kErrNone; k0 += kRun) {
                const uint32_t a = i0 + k0, b = std::min(a + kRun, i1);
                if (a < b) e = xor_run<true>(codec, s, bits, cn, a, b, sink, nullptr);
              }
              expect(e == kErrNone, "segment decode error", codec, g, chunk, k, i0);
            }
            for (uint32_t i = 0; i < cn; ++i)
              expect(as_u64(out[i]) == as_u64(want[i]), "segment value", codec, g, chunk, k, i);
          } else {
            uint32_t e = chimp128_index(data, bits, cn, ck.data());
            expect(e == kErrNone, "Chimp128 index error", codec, g, chunk, k, 0);
            std::vector<uint8_t> dist((cn + kTile - 1) / kTile * kTile + 16, 0);
            C128Term term{};
            for (uint32_t seg = 0; seg < nseg; ++seg) {
              const uint32_t i0 = seg * kSeg, i1 = std::min(i0 + kSeg, cn + 1);
              BitWindow r;
              r.init(data, ck[seg].bitpos);
              uint32_t lead = ck[seg].a == 255 ? 0xffffffffu : ck[seg].a;
              for (uint32_t k0 = 0; k0 < kSeg && e == kErrNone; k0 += kRun) {
                const uint32_t a = i0 + k0, b = std::min(a + kRun, i1);
                uint32_t d16[4] = {0, 0, 0, 0};
                if (a < b) e = chimp128_parse_run(r, lead, bits, cn, a, b, sink, d16, &term);
                if (a < cn) std::memcpy(dist.data() + a, d16, 16);
              }
              expect(e == kErrNone, "Chimp128 parse error", codec, g, chunk, k, i0);
            }
            std::vector<uint64_t> val(cn + 1, 0);
            for (uint32_t i = 0; i <= cn; ++i) {
              const uint64_t x = i < cn ? as_u64(out[i]) : term.x;
              const uint32_t d = i < cn ? dist[i] : term.dist;
              val[i] = (d ? val[i - d] : 0) ^ x;
            }
            for (uint32_t i = 0; i < cn; ++i)
              expect(val[i] == as_u64(want[i]), "Chimp128 value", codec, g, chunk, k, i);
            expect(val[cn] == kNan, "Chimp128 terminator", codec, g, chunk, k, cn);
          }
        }
        ebi_buffer_free(file);
      }
    }
  std::printf("%s: %ld checks, %ld failures\n", fails ? "FAIL" : "OK", checks, fails);
  return fails ? 1 : 0;
}
