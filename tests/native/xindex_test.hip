// Host run of the parallel XOR index (kernels/xindex.hpp) against the serial
// index pass: every chunk of Gorilla / Chimp / Chimp128 / Elf / ElfOnChimp
// files from this library's encoder is indexed both ways -- the serial pass
// (xor_run / chimp128_index writing checkpoints) and the three steps of the
// parallel one (candidates, the chunk walker, emit; the walker one span at
// a time where the kernel takes 64) -- and the checkpoints, the first error
// and the NaN end must agree bit for bit.  Inputs: the four generators at
// ragged chunkings, NaNs at chosen records, and damaged streams (flipped
// bits, a stream cut short).  Also prints the share of spans (of chunks
// with 64 or more) whose candidate missed, which the walker parsed itself.
// Exit code 0 on success.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ebi_hip.h"
#include "../../ebi_amd/csrc/kernels/xindex.hpp"

using namespace ebi;

static long fails = 0, checks = 0;
static uint64_t spans_total = 0, walker_parsed = 0;  // spans the chunk walker parsed itself

static void expect(bool ok, const char *what, int codec, const char *tag, uint64_t chunk, uint64_t k, uint64_t i) {
  ++checks;
  if (!ok && fails++ < 30)
    std::printf("FAIL %s: codec %d %s chunk option %llu, chunk %llu, at %llu\n", what, codec, tag,
                (unsigned long long)chunk, (unsigned long long)k, (unsigned long long)i);
}

struct Event {
  uint32_t code;  // 0 none, kNanEnd, or an error
  uint32_t rec;   // NaN end record (only meaningful for kNanEnd)
};

// The parallel index of one chunk, step by step as the kernels do it:
// candidates (all spans, or Gorilla's windows after each walker stop), the
// chunk walker, then emit.
static Event parallel_index(uint32_t codec, const uint8_t *data, uint32_t bits, uint32_t n, XorCkpt *ck) {
  const uint32_t ns = span_count(bits);
  std::vector<uint64_t> cand(ns, kEntGuess);
  std::vector<SpanSum> res(ns);
  std::vector<SpanPick> pick(ns);
  ChunkWalk w{ent_start(codec), 0, 0, 0, INT32_MAX, 0};
  const bool gorilla = codec == kCodecGorilla;
  if (gorilla)
    chunk_walk_any<BitWindow>(codec, data, bits, ns, w, cand.data(), res.data(), pick.data(), kHintSpans, false);
  for (uint32_t j = 0; j < ns; j += kSpansPerLane) {
    const uint32_t j0 = gorilla ? std::max(j, w.frontier) : j, j1 = std::min(j + kSpansPerLane, ns);
    if (j0 < j1) span_cands_any<BitWindow>(codec, data, bits, ns, j0, j1, gorilla ? w.cur : 0, cand.data(), res.data());
  }
  chunk_walk_any<BitWindow>(codec, data, bits, ns, w, cand.data(), res.data(), pick.data(), ns, gorilla);
  if (ns >= 64) {  // statistics of chunks with many spans
    uint32_t miss = 0;
    for (uint32_t i = 0; i < ns; ++i) miss += pick[i].ent != kEntDead && cand[i] != pick[i].ent;
    spans_total += ns;
    walker_parsed += miss;
  }
  uint64_t first = ~0ull;
  for (uint32_t i = 0; i < ns; ++i) {
    const SpanPick &pk = pick[i];
    const uint64_t ev = span_emit_any<BitWindow>(codec, data, pk.ent, pk.g0, pk.v0, pk.beta0, span_end(i, bits),
                                                 i + 1 == ns, bits, n, ck);
    if (ev < first) first = ev;
  }
  if (first == ~0ull) return Event{0, 0};
  return Event{(uint32_t)(first & 0xff), (uint32_t)(first >> 8)};
}

static Event serial_index(uint32_t codec, const uint8_t *data, uint32_t bits, uint32_t n, XorCkpt *ck) {
  if (codec == kCodecChimp128) return Event{chimp128_index(data, bits, n, ck), 0};
  SerialState s;
  xor_state_start(s, codec, data);
  const DirectSink sink{nullptr};
  const uint32_t e = xor_run<false>(codec, s, bits, n, 0, n, sink, ck);
  return Event{e, e == kNanEnd ? s.end : 0};
}

static void compare_chunk(uint32_t codec, const uint8_t *data, uint32_t bits, uint32_t n, const char *tag,
                          uint64_t chunk, uint64_t k) {
  const uint32_t every = codec == kCodecChimp128 ? kSeg128 : kSeg;
  const uint32_t nck = n / every + 2;
  std::vector<XorCkpt> a(nck), b(nck);
  std::memset(a.data(), 0xcd, nck * sizeof(XorCkpt));
  std::memset(b.data(), 0xcd, nck * sizeof(XorCkpt));
  const Event es = serial_index(codec, data, bits, n, a.data());
  const Event ep = parallel_index(codec, data, bits, n, b.data());
  expect(es.code == ep.code, "event code", codec, tag, chunk, k, es.code * 1000 + ep.code);
  if (es.code == kNanEnd && ep.code == kNanEnd) expect(es.rec == ep.rec, "NaN end record", codec, tag, chunk, k, es.rec);
  // checkpoints the serial pass wrote (it stops at its first event)
  for (uint32_t j = 0; j < nck; ++j) {
    XorCkpt cc;
    std::memset(&cc, 0xcd, sizeof cc);
    if (std::memcmp(&a[j], &cc, sizeof cc) == 0) continue;
    expect(std::memcmp(&a[j], &b[j], sizeof(XorCkpt)) == 0, "checkpoint", codec, tag, chunk, k, j);
  }
}

static void run_file(int codec, const double *v, uint64_t n, uint64_t chunk, const char *tag, bool damage) {
  ebi_encoder_config cfg{};
  cfg.scheme = codec;
  cfg.chunk_kind = EBI_CHUNK_RECORD_COUNT;
  cfg.chunk_value = chunk;
  uint8_t *file = nullptr;
  size_t len = 0;
  if (ebi_encode(v, n, &cfg, &file, &len) != EBI_OK) {
    ++fails;
    return;
  }
  ebi_file_info info{};
  ebi_file_parse(file, len, &info);
  const size_t nc = info.number_of_chunks;
  std::vector<uint64_t> off(nc), size(nc), loff(nc), nrec(nc);
  ebi_file_chunks(file, len, off.data(), size.data(), loff.data(), nrec.data());
  for (size_t k = 0; k < nc; ++k) {
    const uint32_t cn = (uint32_t)nrec[k], bits = (uint32_t)(size[k] * 8);
    std::vector<uint4> body((size[k] + 256 + 15) / 16, make_uint4(0, 0, 0, 0));
    std::memcpy(body.data(), file + off[k], size[k]);
    uint8_t *data = reinterpret_cast<uint8_t *>(body.data());
    compare_chunk(codec, data, bits, cn, tag, chunk, k);
    if (damage && bits > 64) {
      // a flipped bit (the body's padding stays zero), then the same stream
      // declared 40 bits shorter
      std::vector<uint4> bad(body);
      uint8_t *bd = reinterpret_cast<uint8_t *>(bad.data());
      for (int t = 0; t < 3; ++t) {
        const uint32_t at = (uint32_t)(((k + 1) * 2654435761u + chunk * 40503u + t * 7919u) % bits);
        std::vector<uint4> one(body);
        uint8_t *od = reinterpret_cast<uint8_t *>(one.data());
        od[at / 8] ^= (uint8_t)(0x80u >> (at % 8));
        compare_chunk(codec, od, bits, cn, "flip", chunk, k);
      }
      (void)bd;
      compare_chunk(codec, data, bits > 40 ? bits - 40 : 0, cn, "cut", chunk, k);
    }
  }
  ebi_buffer_free(file);
}

int main() {
  struct Gen {
    int kind, dec;
    double p0, p1;
    const char *name;
  } gens[] = {{0, 2, 100.0, 0.1, "walk2"}, {1, 3, -1000.0, 1000.0, "unif3"}, {2, 2, 0.01, 0.0, "step2"},
              {3, 4, 1000.0, 0.01, "sin4"}};
  const int codecs[] = {kCodecGorilla, kCodecChimp, kCodecElf, kCodecElfOnChimp, kCodecChimp128};
  const uint64_t chunks[] = {1, 3, 64, 65, 1003, 4096, 65536};
  const uint64_t n = 140003;
  std::vector<double> v(n);
  for (int codec : codecs) {
    for (const Gen &g : gens) {
      walker_parsed = spans_total = 0;
      if (ebi_synth_column(g.kind, n, 11, g.dec, g.p0, g.p1, 0.0, 1, v.data()) != EBI_OK) return 2;
      for (uint64_t chunk : chunks) {
        // short chunkings on a prefix (one record per chunk is a chunk per record)
        const uint64_t nn = chunk < 64 ? 3001 : n;
        run_file(codec, v.data(), nn, chunk, g.name, chunk == 1003 || chunk == 4096);
      }
      std::printf("codec %d %s: the walker parsed %llu of %llu spans\n", codec, g.name,
                  (unsigned long long)walker_parsed, (unsigned long long)spans_total);
    }
    // NaNs: record 0, mid-chunk, a chunk's last record, two in one chunk,
    // right after a checkpoint, the column's last record
    const uint64_t m = 20000;
    if (ebi_synth_column(0, m, 99, 2, 100.0, 0.1, 0.0, 1, v.data()) != EBI_OK) return 2;
    for (uint64_t i : {0ull, 1500ull, 2005ull, 3100ull, 3200ull, 3328ull, 9000ull, 19999ull}) v[i] = as_f64(kNan);
    for (uint64_t chunk : {1003ull, 64ull, 65ull, 8192ull}) run_file(codec, v.data(), m, chunk, "nan", false);
  }
  std::printf("%s: %ld checks, %ld failures\n", fails ? "FAIL" : "OK", checks, fails);
  return fails ? 1 : 0;
}
