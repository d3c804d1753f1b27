// Host run of the XOR decoders (Gorilla, Chimp, Elf, ElfOnChimp, Chimp128) on files
// from this library's encoder.  Every Gorilla / Chimp / Elf chunk is decoded
// whole from the stream start, and in the segment kernel's forms: the index
// pass writes a checkpoint every kSeg records, then each segment restarts
// from its checkpoint and decodes in runs of kRun records with the state
// carried between runs -- per record, and in the kernel's lockstep form
// (every lane of a wave through the same record loop) through the column's
// lane-interleaved copy, with a check that no lane reads past the blocks the
// copy gives it; Gorilla and Chimp also on damaged streams (both forms stop
// each segment with the same error).  Chimp128 as the segment kernel's job
// does it: a checkpoint every kSeg128 records, the lockstep parse of 64
// segments, then the values in record order from the job's ring checkpoint.
// All must give back the encoded values bit for bit (the codecs are
// lossless).  Synthetic columns of the four generators, ragged chunkings
// down to one record per chunk.  Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ebi_hip.h"
#include "../../ebi_amd/csrc/kernels/codecs.hpp"

using namespace ebi;

static long fails = 0, checks = 0;
static double g_elf_tab[kElfTab];  // the segment kernel's LDS table of powers of ten

static void expect(bool ok, const char *what, int codec, int gen, uint64_t chunk, uint64_t k, uint64_t i) {
  ++checks;
  if (!ok && fails++ < 20)
    std::printf("FAIL %s: codec %d gen %d chunk option %llu, chunk %llu, record %llu\n", what, codec, gen,
                (unsigned long long)chunk, (unsigned long long)k, (unsigned long long)i);
}

// The column's lane-interleaved segment streams as the load-time kernels
// build them (same block counts, ilv_lane_blocks), and the check that a
// lane's window never requested a block past its own: after a decode the
// window's next request is at p, so (p - lane_base) / 64 blocks were taken.
struct Ilv {
  std::vector<uint4> blocks;
  std::vector<uint32_t> job_row, nb;  // per job: first row; per segment: block count
  const uint4 *lane(uint32_t seg) const {
    return blocks.data() + (size_t)job_row[seg / kWave] * kIlvRowBlocks + seg % kWave;
  }
};

static Ilv build_ilv(const uint8_t *data, uint32_t bits, const XorCkpt *ck, uint32_t nseg) {
  Ilv v;
  const uint4 *src = reinterpret_cast<const uint4 *>(data);
  v.nb.resize(nseg);
  for (uint32_t first = 0; first < nseg; first += kWave) {
    uint32_t rows = 0;
    for (uint32_t seg = first; seg < std::min(nseg, first + kWave); ++seg) {
      const uint32_t a = std::min(ck[seg].bitpos, bits);
      const uint32_t b = seg + 1 < nseg ? std::min(std::max(ck[seg + 1].bitpos, a), bits) : bits;
      v.nb[seg] = ilv_lane_blocks(a, b);
      rows = std::max(rows, v.nb[seg]);
    }
    const uint32_t row0 = (uint32_t)(v.blocks.size() / kIlvRowBlocks);
    v.job_row.push_back(row0);
    v.blocks.resize(v.blocks.size() + (size_t)rows * kIlvRowBlocks, make_uint4(0, 0, 0, 0));
    for (uint32_t seg = first; seg < std::min(nseg, first + kWave); ++seg)
      for (uint32_t r = 0; r < v.nb[seg]; ++r)
        v.blocks[((size_t)row0 + r) * kIlvRowBlocks + seg % kWave] = src[(ck[seg].bitpos >> 7) + r];
  }
  return v;
}

static bool within_lane(const Ilv &v, uint32_t seg, const BitWindow &r) {
  return (size_t)(r.p - v.lane(seg)) <= (size_t)v.nb[seg] * kIlvRowBlocks;
}

// The segment kernel's lockstep form, wave by wave: 64 segments per wave
// job, lanes without a segment included (their window has no stream, so a
// read through it would crash this host run), runs of kRun records with the
// wave's shared trip count.
struct RowSink {
  double *out;
  uint32_t lim;  // records the buffer holds
  void row(uint32_t i, uint32_t k, double v, bool act = true) const {  // row kRun: the stage row never flushed
    if (act && k < kRun && i < lim) out[i] = v;
  }
};

// One lockstep run of a Gorilla / Chimp / Elf / ElfOnChimp lane (with the
// per-record checks, or without them as the segment kernels run accepted
// chunks), and a lane's record 0.
template <bool kChecked, class Sink>
static uint32_t seg_step(int codec, XorState &s, uint32_t bits, uint32_t n, uint32_t a, uint32_t cnt, uint32_t from,
                         uint32_t lim, Sink &sink, uint32_t e) {
  switch (codec) {
    case kCodecGorilla: return gorilla_seg<kChecked>(s, bits, n, a, cnt, from, lim, sink, e);
    case kCodecChimp: return chimp_seg<kChecked>(s, bits, n, a, cnt, from, lim, sink, e);
    case kCodecElfOnChimp: return elfchimp_seg<kChecked>(s, bits, n, a, cnt, from, lim, sink, e, g_elf_tab);
    default: return elf_seg<kChecked>(s, bits, n, a, cnt, from, lim, sink, e, g_elf_tab);
  }
}
template <class Sink>
static uint32_t seg_first(int codec, XorState &s, uint32_t bits, Sink &sink) {
  switch (codec) {
    case kCodecElf: return elf_seg_first(s, bits, sink, g_elf_tab);
    case kCodecElfOnChimp: return elfchimp_seg_first(s, bits, sink, g_elf_tab);
    default: return xor_seg_first(s, bits, sink);
  }
}

template <bool kChecked = true>
static std::vector<uint32_t> lockstep(int codec, const uint8_t *data, uint32_t bits, uint32_t cn, const XorCkpt *ck,
                                      uint32_t nseg, double *out, const Ilv *ilv = nullptr) {
  RowSink sink{out, cn};
  std::vector<uint32_t> err(nseg, 0);
  for (uint32_t first = 0; first < nseg; first += kWave) {
    XorState st[kWave];
    uint32_t e[kWave], from[kWave], last[kWave], i0[kWave];
    for (uint32_t l = 0; l < kWave; ++l) {
      i0[l] = (first + l) * kSeg;
      const bool live = i0[l] < cn;
      const uint32_t i1 = live ? std::min(i0[l] + kSeg, cn) : 0;
      if (live && ilv) {
        st[l].r.init_at(ilv->lane(first + l), ck[first + l].bitpos, kIlvRowBlocks);
        xor_state_set(st[l], codec, ck[first + l]);
      } else if (live) {
        xor_state_at(st[l], codec, data, ck[first + l]);
      } else {
        st[l] = XorState{};
        st[l].r.p = nullptr;
      }
      last[l] = live && i1 == cn ? cn + 1 : i1;
      from[l] = i0[l];
      e[l] = 0;
      if (live && i0[l] == 0) {
        e[l] = seg_first(codec, st[l], bits, sink);
        from[l] = 1;
      }
    }
    for (uint32_t k0 = 0; k0 < kSeg; k0 += kRun) {
      bool extra = false;
      for (uint32_t l = 0; l < kWave; ++l) extra |= last[l] == i0[l] + k0 + kRun + 1;
      const uint32_t cnt = kRun + (extra ? 1 : 0);
      for (uint32_t l = 0; l < kWave; ++l) {
        const uint32_t a = i0[l] + k0;
        const uint32_t lim = last[l] == a + kRun + 1 ? last[l] : std::min(last[l], a + kRun);
        e[l] = seg_step<kChecked>(codec, st[l], bits, cn, a, cnt, from[l], lim, sink, e[l]);
        from[l] = std::max(from[l], lim);
      }
    }
    for (uint32_t l = 0; l < kWave && first + l < nseg; ++l) {
      err[first + l] = e[l];
      if (ilv && !within_lane(*ilv, first + l, st[l].r)) err[first + l] |= 0x100;  // read past the lane's blocks
    }
  }
  return err;
}

// The per-record form over the same segments (the reference's order of
// checks); per-segment errors.
static std::vector<uint32_t> segform(int codec, const uint8_t *data, uint32_t bits, uint32_t cn, const XorCkpt *ck,
                                     uint32_t nseg, double *out) {
  const DirectSink sink{out};
  std::vector<uint32_t> err(nseg, 0);
  XorState s;
  for (uint32_t seg = 0; seg < nseg; ++seg) {
    const uint32_t i0 = seg * kSeg, i1 = std::min(i0 + kSeg, cn);
    xor_state_at(s, codec, data, ck[seg]);
    uint32_t e = 0;
    for (uint32_t k0 = 0; k0 < kSeg && e == kErrNone; k0 += kRun) {
      const uint32_t a = i0 + k0, b = std::min(a + kRun, i1);
      if (a < b) e = xor_run<true>(codec, s, bits, cn, a, b, sink, nullptr);
    }
    // a NaN before the count: the load-time pass would have shortened the
    // chunk there; over the unshortened count the lockstep form reports it
    // as a count error
    err[seg] = e == kNanEnd ? (uint32_t)kErrCount : e;
  }
  return err;
}

// Columns holding canonical NaNs (0x7ff8000000000000, numpy's nan): the
// reference's XOR readers stop a chunk at its first NaN and return the
// records before it (chunk_reader/gorilla.rs:149-152, chimp.rs:119,
// chimp_n.rs:219-265, elf.rs:105-258).  The load-time pass must report the
// NaN's index (kNanEnd, XorState::end) instead of an error, and the segment
// forms, run with the chunk shortened to that count, must decode the records
// before it, with dead lanes whose state is poisoned (only the window's bit
// offset is defined, as a dead lane never advances): no result may depend
// on a dead lane's registers.
static void nan_cases(int codec) {
  const uint64_t n = 5000;
  std::vector<double> v(n);
  if (ebi_synth_column(0, n, 99, 2, 100.0, 0.1, 0.0, 1, v.data()) != EBI_OK) {
    ++fails;
    return;
  }
  const double qnan = as_f64(kNan);
  // record 0 of chunk 0, mid-chunk, a chunk's last record, two NaNs in one
  // chunk, a NaN right after a segment boundary, the column's last record
  for (uint64_t i : {0ull, 1500ull, 2005ull, 3100ull, 3200ull, 3328ull, 4999ull}) v[i] = qnan;
  for (uint64_t chunk : {1003ull, 64ull, 65ull}) {
    ebi_encoder_config cfg{};
    cfg.scheme = codec;
    cfg.chunk_kind = EBI_CHUNK_RECORD_COUNT;
    cfg.chunk_value = chunk;
    uint8_t *file = nullptr;
    size_t len = 0;
    if (ebi_encode(v.data(), n, &cfg, &file, &len) != EBI_OK) {
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
      const uint8_t *data = reinterpret_cast<const uint8_t *>(body.data());
      const double *want = v.data() + loff[k];
      uint32_t end = cn;  // the first NaN of the chunk, if any
      for (uint32_t i = 0; i < cn; ++i)
        if (as_u64(want[i]) == kNan) {
          end = i;
          break;
        }
      std::vector<double> out(cn + 1, 0.0);
      if (codec == kCodecChimp128) {
        // parse-only index over the footer count, then the job resolve of
        // the chunk shortened to `end` records: records below it and the
        // NaN at `end`
        const uint32_t nseg = cn / kSeg128 + 1;
        std::vector<XorCkpt> ck(nseg + 1);
        expect(chimp128_index(data, bits, cn, ck.data()) == kErrNone, "NaN: Chimp128 index", codec, 9, chunk, k, 0);
        const uint32_t nseg_e = end / kSeg128 + 1;
        std::vector<uint64_t> res(end + 1, 0);
        for (uint32_t seg = 0; seg < nseg_e; ++seg) {
          const uint32_t i0 = seg * kSeg128;
          std::vector<double> xs(kSeg128 + 1, 0.0);
          std::vector<uint32_t> ds(kSeg128 + 1, 0);
          struct SegSink {
            double *x;
            uint32_t *d;
            void row(uint32_t, uint32_t kk, double val, bool act) {
              if (act) x[kk] = val;
            }
            void dist(uint32_t kk, uint32_t val, bool act) {
              if (act) d[kk] = val;
            }
          } js{xs.data(), ds.data()};
          BitWindow r;
          r.init(data, ck[seg].bitpos);
          uint32_t lead = ck[seg].a != 255 ? ck[seg].a : 0xffffffffu, from = i0, ee = 0;
          uint64_t px = 0;
          uint32_t pd = 0;
          if (i0 == 0) {
            ee = chimp128_seg_first(r, bits, js, px, pd);
            from = 1;
          }
          ee = chimp128_parse_seg<true>(r, lead, bits, i0, kSeg128, from, std::min(i0 + kSeg128, end + 1), js, ee, px, pd,
                                  i0 == 0);
          expect(ee == kErrNone, "NaN: Chimp128 parse", codec, 9, chunk, k, i0);
          for (uint32_t t = 0; t < kSeg128 && i0 + t <= end; ++t) {
            const uint32_t i = i0 + t, d = ds[t];
            res[i] = as_u64(xs[t]) ^ (d ? res[i - d] : 0ull);
          }
        }
        for (uint32_t i = 0; i < end; ++i)
          expect(res[i] == as_u64(want[i]), "NaN: Chimp128 value", codec, 9, chunk, k, i);
        if (end < cn) expect(res[end] == kNan, "NaN: Chimp128 end marker", codec, 9, chunk, k, end);
        continue;
      }
      // index pass over the footer count: kNanEnd at the first NaN
      std::vector<XorCkpt> ck((cn + kSeg - 1) / kSeg + 1);
      SerialState s{};  // the device index pass's form (deep prefetch window)
      xor_state_start(s, codec, data);
      const DirectSink sink{out.data()};
      const uint32_t e = xor_run<false>(codec, s, bits, cn, 0, cn, sink, ck.data());
      if (end < cn) {
        expect(e == kNanEnd && s.end == end, "NaN: index pass end", codec, 9, chunk, k, end);
      } else {
        expect(e == kErrNone, "NaN: index pass", codec, 9, chunk, k, 0);
      }
      // lockstep over the shortened chunk, dead lanes poisoned
      const uint32_t nseg = (end + kSeg - 1) / kSeg;
      const Ilv ilv = build_ilv(data, bits, ck.data(), nseg);
      RowSink rs{out.data(), end};
      std::fill(out.begin(), out.end(), 0.0);
      for (uint32_t first = 0; first < std::max(nseg, 1u); first += kWave) {
        XorState st[kWave];
        uint32_t ee[kWave], from[kWave], last[kWave], i0[kWave];
        for (uint32_t l = 0; l < kWave; ++l) {
          i0[l] = (first + l) * kSeg;
          const bool live = i0[l] < end;
          const uint32_t i1 = live ? std::min(i0[l] + kSeg, end) : 0;
          if (live) {
            st[l].r.init_at(ilv.lane(first + l), ck[first + l].bitpos, kIlvRowBlocks);
            xor_state_set(st[l], codec, ck[first + l]);
          } else {
            std::memset(static_cast<void *>(&st[l]), 0xa5, sizeof(XorState));
            st[l].r.b = 0;
          }
          last[l] = live && i1 == end ? end + 1 : i1;
          from[l] = i0[l];
          ee[l] = 0;
          if (live && i0[l] == 0) {
            ee[l] = seg_first(codec, st[l], bits, rs);
            from[l] = 1;
          }
        }
        for (uint32_t k0 = 0; k0 < kSeg; k0 += kRun) {
          bool extra = false;
          for (uint32_t l = 0; l < kWave; ++l) extra |= last[l] == i0[l] + k0 + kRun + 1;
          const uint32_t cnt = kRun + (extra ? 1 : 0);
          for (uint32_t l = 0; l < kWave; ++l) {
            const uint32_t a = i0[l] + k0;
            const uint32_t lim = last[l] == a + kRun + 1 ? last[l] : std::min(last[l], a + kRun);
            ee[l] = seg_step<true>(codec, st[l], bits, end, a, cnt, from[l], lim, rs, ee[l]);
            from[l] = std::max(from[l], lim);
          }
        }
        for (uint32_t l = 0; l < kWave; ++l)
          if (i0[l] < end) expect(ee[l] == kErrNone, "NaN: lockstep error", codec, 9, chunk, k, i0[l]);
      }
      for (uint32_t i = 0; i < end; ++i)
        expect(as_u64(out[i]) == as_u64(want[i]), "NaN: lockstep value", codec, 9, chunk, k, i);
    }
    ebi_buffer_free(file);
  }
}

int main() {
  char buf[32];
  for (int i = 0; i < 400; ++i) {  // the host copies of the tables, filled as the library fills the device's
    std::snprintf(buf, sizeof buf, "1.0E%d", i);
    c_p10[i] = std::strtod(buf, nullptr);
    std::snprintf(buf, sizeof buf, "1.0E-%d", i);
    c_n10[i] = std::strtod(buf, nullptr);
  }
  for (uint32_t i = 0; i < kElfTab; ++i) elf_tab_fill(g_elf_tab, i, c_p10, c_n10);
  struct Gen {
    int kind, dec;
    double p0, p1;
  } gens[] = {{0, 2, 100.0, 0.1}, {1, 3, -1000.0, 1000.0}, {2, 2, 0.01, 0.0}, {3, 4, 1000.0, 0.01}};
  const int codecs[] = {kCodecGorilla, kCodecChimp, kCodecElf, kCodecElfOnChimp, kCodecChimp128};
  const uint64_t chunks[] = {1, 3, 5, 64, 65, 80, 96, 112, 127, 1003, 4096};  // 80 / 96 / 112: terminator after a full run
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
          const uint32_t nseg = codec == kCodecChimp128 ? cn / kSeg128 + 1 : (cn + kSeg - 1) / kSeg;
          std::vector<XorCkpt> ck(nseg + 1);
          XorState s;
          if (codec != kCodecChimp128) {
            // whole chunk and index pass in the lane-per-chunk form (deep
            // prefetch window), both windows over the same padded body
            SerialState ss;
            xor_state_start(ss, codec, data);
            uint32_t e = xor_run<true>(codec, ss, bits, cn, 0, cn, sink, nullptr);
            expect(e == kErrNone, "whole-chunk decode error", codec, g, chunk, k, 0);
            for (uint32_t i = 0; i < cn; ++i)
              expect(as_u64(out[i]) == as_u64(want[i]), "whole-chunk value", codec, g, chunk, k, i);
            xor_state_start(s, codec, data);
            e = xor_run<true>(codec, s, bits, cn, 0, cn, sink, nullptr);
            expect(e == kErrNone, "whole-chunk decode error (depth 1)", codec, g, chunk, k, 0);
            xor_state_start(ss, codec, data);
            e = xor_run<false>(codec, ss, bits, cn, 0, cn, sink, ck.data());
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
            const Ilv ilv = build_ilv(data, bits, ck.data(), nseg);
            if (codec == kCodecElf || codec == kCodecElfOnChimp) {  // the per-record form through the ilv
              std::fill(out.begin(), out.end(), 0.0);
              for (uint32_t seg = 0; seg < nseg; ++seg) {
                const uint32_t i0 = seg * kSeg, i1 = std::min(i0 + kSeg, cn);
                s.r.init_at(ilv.lane(seg), ck[seg].bitpos, kIlvRowBlocks);
                xor_state_set(s, codec, ck[seg]);
                for (uint32_t k0 = 0; k0 < kSeg && e == kErrNone; k0 += kRun) {
                  const uint32_t a = i0 + k0, b = std::min(a + kRun, i1);
                  if (a < b) e = xor_run<true>(codec, s, bits, cn, a, b, sink, nullptr);
                }
                expect(e == kErrNone, "ilv segment decode error", codec, g, chunk, k, i0);
                expect(within_lane(ilv, seg, s.r), "ilv lane overread", codec, g, chunk, k, i0);
              }
              for (uint32_t i = 0; i < cn; ++i)
                expect(as_u64(out[i]) == as_u64(want[i]), "ilv segment value", codec, g, chunk, k, i);
            }
            {
              std::fill(out.begin(), out.end(), 0.0);
              const std::vector<uint32_t> le = lockstep(codec, data, bits, cn, ck.data(), nseg, out.data(), &ilv);
              for (uint32_t sg = 0; sg < nseg; ++sg)
                expect(le[sg] == kErrNone, "lockstep (ilv) error", codec, g, chunk, k, sg * kSeg);
              for (uint32_t i = 0; i < cn; ++i)
                expect(as_u64(out[i]) == as_u64(want[i]), "lockstep (ilv) value", codec, g, chunk, k, i);
              // the kernels' form: no per-record checks on the accepted chunk
              std::fill(out.begin(), out.end(), 0.0);
              const std::vector<uint32_t> lu = lockstep<false>(codec, data, bits, cn, ck.data(), nseg, out.data(), &ilv);
              for (uint32_t sg = 0; sg < nseg; ++sg)
                expect(lu[sg] == kErrNone, "unchecked lockstep error", codec, g, chunk, k, sg * kSeg);
              for (uint32_t i = 0; i < cn; ++i)
                expect(as_u64(out[i]) == as_u64(want[i]), "unchecked lockstep value", codec, g, chunk, k, i);
              // damaged streams (checkpoints of the intact one): a flipped
              // bit, then a stream cut short; both forms must stop each
              // segment with the same error, and agree on segments that
              // decode
              std::vector<uint4> bad(body);
              uint8_t *bd = reinterpret_cast<uint8_t *>(bad.data());
              std::vector<double> o1(cn + 1), o2(cn + 1);
              for (int trial = 0; trial < 2; ++trial) {
                uint32_t bb = bits;
                if (trial == 0) {
                  const uint32_t at = (uint32_t)((k * 2654435761u + chunk) % bits);
                  bd[at / 8] ^= (uint8_t)(0x80u >> (at % 8));
                } else {
                  bd = const_cast<uint8_t *>(data);
                  bb = bits > 40 ? bits - 40 : 0;
                }
                const std::vector<uint32_t> e1 = segform(codec, bd, bb, cn, ck.data(), nseg, o1.data());
                const std::vector<uint32_t> e2 = lockstep(codec, bd, bb, cn, ck.data(), nseg, o2.data());
                for (uint32_t sg = 0; sg < nseg; ++sg) {
                  expect(e1[sg] == e2[sg], "damaged stream: error", codec, g, chunk, k, sg * kSeg);
                  if (e1[sg] == kErrNone)
                    for (uint32_t i = sg * kSeg; i < std::min(cn, sg * kSeg + kSeg); ++i)
                      expect(as_u64(o1[i]) == as_u64(o2[i]), "damaged stream: value", codec, g, chunk, k, i);
                }
              }
            }
          } else {
            uint32_t e = chimp128_index(data, bits, cn, ck.data());
            expect(e == kErrNone, "Chimp128 index error", codec, g, chunk, k, 0);
            // the segment kernel's Chimp128 job, wave by wave: the lockstep
            // parse of 64 segments of kSeg128 records (in place), then the
            // values in record order from the job's ring checkpoint (the
            // 128 values before it)
            for (uint32_t first = 0; first < nseg; first += kWave) {
              const uint32_t j0 = first * kSeg128;
              std::vector<double> xs(kJob128, 0.0);
              std::vector<uint32_t> ds(kJob128, 0);
              for (uint32_t l = 0; l < kWave; ++l) {
                const uint32_t seg = first + l, i0 = seg * kSeg128;
                if (seg >= nseg) break;
                struct JobSink {
                  double *x;
                  uint32_t *d;
                  void row(uint32_t, uint32_t kk, double v, bool act) {
                    if (act) x[kk] = v;
                  }
                  void dist(uint32_t kk, uint32_t v, bool act) {
                    if (act) d[kk] = v;
                  }
                } js{xs.data() + l * kSeg128, ds.data() + l * kSeg128};
                BitWindow r;
                r.init(data, ck[seg].bitpos);
                uint32_t lead = ck[seg].a != 255 ? ck[seg].a : 0xffffffffu, from = i0, ee = 0;
                uint64_t px = 0;
                uint32_t pd = 0;
                if (i0 == 0) {
                  ee = chimp128_seg_first(r, bits, js, px, pd);
                  from = 1;
                }
                // the kernels' form (no per-record checks) from the same state
                BitWindow ru = r;
                uint32_t leadu = lead;
                std::vector<double> xu(kSeg128, 0.0);
                std::vector<uint32_t> du(kSeg128, 0);
                JobSink ju{xu.data(), du.data()};
                const uint32_t eu = chimp128_parse_seg<false>(ru, leadu, bits, i0, kSeg128, from,
                                                              std::min(i0 + kSeg128, cn + 1), ju, ee, px, pd, i0 == 0);
                ee = chimp128_parse_seg<true>(r, lead, bits, i0, kSeg128, from, std::min(i0 + kSeg128, cn + 1), js, ee,
                                              px, pd, i0 == 0);
                expect(ee == kErrNone, "Chimp128 parse error", codec, g, chunk, k, i0);
                expect(eu == kErrNone, "unchecked Chimp128 parse error", codec, g, chunk, k, i0);
                for (uint32_t t = from - i0; t < kSeg128 && i0 + t <= cn; ++t)  // record 0: parsed before
                  expect(as_u64(xu[t]) == as_u64(xs[l * kSeg128 + t]) && du[t] == ds[l * kSeg128 + t],
                         "unchecked Chimp128 parse", codec, g, chunk, k, i0 + t);
              }
              std::vector<uint64_t> res(kJob128, 0);
              for (uint32_t t = 0; t < kJob128 && j0 + t <= cn; ++t) {
                const uint32_t i = j0 + t, d = ds[t];
                uint64_t src = 0;
                if (d && i - d < j0)  // before the job: the ring checkpoint holds only the 128 records before it
                  expect(i - d + 128 >= j0, "Chimp128 source before the ring checkpoint", codec, g, chunk, k, i);
                if (d) src = i - d >= j0 ? res[i - d - j0] : as_u64(want[i - d]);
                res[t] = as_u64(xs[t]) ^ src;
                if (i < cn)
                  expect(res[t] == as_u64(want[i]), "Chimp128 value", codec, g, chunk, k, i);
                else
                  expect(res[t] == kNan, "Chimp128 terminator", codec, g, chunk, k, cn);
              }
            }
          }
        }
        ebi_buffer_free(file);
      }
    }
  for (int codec : codecs) nan_cases(codec);
  std::printf("%s: %ld checks, %ld failures\n", fails ? "FAIL" : "OK", checks, fails);
  return fails ? 1 : 0;
}
