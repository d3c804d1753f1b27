// Per-codec device decoders for libebi_hip (gfx950).
//
// Two shapes of work:
//  * tile codecs (Uncompressed, BUFF, DeltaSprintz, RLE): one wave decodes a
//    tile of kTile = 512 records (fields for RLE), 8 consecutive records per
//    lane, with wave scans for the prefix-dependent codecs; the prefix state
//    at each tile start comes from the column's tile index;
//  * XOR codecs (Gorilla, Chimp, Chimp128, Elf): the bit length of a record
//    depends on state set by earlier records, so one lane walks one chunk
//    (or one kSeg-record segment, restarting from a checkpoint written by the
//    load-time index pass) serially.
// All functions restate the reference decoders cited beside them
// (paths relative to lemolatoon/ebi).
#pragma once

#include "common.hpp"

namespace ebi {

// Powers of ten as the reference obtains them (compression_common/elf.rs:95-109):
// correctly rounded parses of "1.0E{i}" / "1.0E-{i}"; filled by the host.
__constant__ double c_p10[400];
__constant__ double c_n10[400];

__host__ __device__ __forceinline__ int32_t sat_i32(double d) {
  if (d != d) return 0;
  if (d >= 2147483647.0) return INT32_MAX;
  if (d <= -2147483648.0) return INT32_MIN;
  return (int32_t)d;
}

// get_sp, compression_common/elf.rs:111-130.  The reference walks a table
// for v in [1e-10, 1e9) and takes floor(log10(v)) otherwise; the walk is
// restated as counts of threshold compares (no divergent loop): for v >= 1
// it returns the first i in 0..8 with v < 1e{i+1}, i.e. #{j in 1..9 : v >=
// 1e{j}}; for v < 1 the first i in 1..10 with v >= 1e-{i}, i.e. 11 - #{j in
// 1..10 : v >= 1e-{j}} (a count of 0 -- v below 1e-10, zero, NaN -- falls
// through).  The literals are the correctly rounded values the reference
// parses from "1.0E{i}" / "1.0E-{i}".
__host__ __device__ __forceinline__ int32_t elf_get_sp(double v) {
  if (v >= 1.0) {
    const int32_t c = (v >= 1e1) + (v >= 1e2) + (v >= 1e3) + (v >= 1e4) + (v >= 1e5) + (v >= 1e6) + (v >= 1e7) +
                      (v >= 1e8) + (v >= 1e9);
    if (c < 9) return c;
  } else {
    const int32_t d = (v >= 1e-1) + (v >= 1e-2) + (v >= 1e-3) + (v >= 1e-4) + (v >= 1e-5) + (v >= 1e-6) +
                      (v >= 1e-7) + (v >= 1e-8) + (v >= 1e-9) + (v >= 1e-10);
    if (d > 0) return d - 11;
  }
  return sat_i32(floor(log10(v)));
}

// 10^a for a <= 22, exactly representable: each partial product below is an
// exact power of ten, so the result equals the reference's parse of "1.0E{a}".
__host__ __device__ __forceinline__ double p10_exact(uint32_t a) {
  double s = (a & 1) ? 10.0 : 1.0;
  s *= (a & 2) ? 100.0 : 1.0;
  s *= (a & 4) ? 1e4 : 1.0;
  s *= (a & 8) ? 1e8 : 1.0;
  s *= (a & 16) ? 1e16 : 1.0;
  return s;
}

// round_up, compression_common/elf.rs:51-58 (no FMA contraction).
__host__ __device__ __forceinline__ double elf_round_up(double v, uint32_t alpha) {
  const double scale = alpha <= 22 ? p10_exact(alpha)
                       : alpha < 400 ? c_p10[alpha]
                                     : as_f64(0x7ff0000000000000ull);
  const double t = v * scale;
  return (v < 0.0 ? floor(t) : ceil(t)) / scale;
}

// recover_v_by_beta_star, decoder/chunk_reader/elf.rs:66-81: the value of
// an erased record from its stored bits v' and beta*.
__host__ __device__ __forceinline__ double elf_recover(double v, int32_t beta) {
  const int32_t sp = elf_get_sp(fabs(v));
  if (beta == 0) {
    const uint32_t k = 0u - (uint32_t)sp - 1u;  // (-sp - 1) as u32, wrapping
    const double w = k < 400 ? c_n10[k] : 0.0;
    return v < 0.0 ? -w : w;
  }
  return elf_round_up(v, (uint32_t)beta - (uint32_t)sp - 1u);
}

// elf_recover with the powers of ten in a small table, for the segment
// kernel (which keeps it in LDS).  tab[k + 10] = 1e{k} for k in -10..9
// (get_sp's thresholds), tab[20 + a] = 1e{a} and tab[43 + a] = 1e-{a} for
// a in 0..22, all the reference's parses of "1.0E{k}" (elf_tab_fill).
// Results equal elf_recover's for every input:
//  * get_sp: for x = |v| in [1e-10, 1e9) the table walk returns the largest
//    k with x >= 1e{k}.  x in [2^e, 2^{e+1}) puts log10(x) in
//    [e log10(2), (e + 1) log10(2)), an interval shorter than 1, so that k
//    is k0 = floor(e log10(2)) or k0 + 1: one compare with 1e{k0 + 1}.
//    floor(e log10(2)) = (e * 1233) >> 12 for the e of that range;
//  * round_up's division N / D, D = 10^alpha exact (alpha <= 22), N a
//    nonzero integer-valued double below 1e31 in magnitude: with y = 1e-alpha
//    = RN(1/D), q0 = RN(N y) is within 1.5 ulp of N / D, one Newton
//    correction q1 = RN(q0 + RN(N - q0 D) y) within one ulp, and the second,
//    q2 = RN(q1 + r1 y) with r1 = N - q1 D exact by FMA, is RN(N / D) by
//    Markstein's theorem (P. Markstein, "Computation of elementary functions
//    on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990).
// Other inputs (x outside [1e-10, 1e9), alpha > 22, beta* = 0 with a power
// past 1e-22) take elf_recover's path.
constexpr uint32_t kElfTab = 66;

// RN(N / D) for D >= 1 and N integer-valued doubles (N = 0 or |N| >= 1, so
// the quotient is zero or normal) and y = RN(1 / D): the product N y and two
// FMA corrections (Markstein; see above).  Used for the Elf powers of ten
// and the DeltaSprintz scale.
__host__ __device__ __forceinline__ double div_rn(double N, double D, double y) {
  const double q0 = N * y;
  const double q1 = fma(fma(-q0, D, N), y, q0);
  return fma(fma(-q1, D, N), y, q1);
}

__host__ __device__ __forceinline__ void elf_tab_fill(double *tab, uint32_t i, const double *p10, const double *n10) {
  if (i < 20) tab[i] = i >= 10 ? p10[i - 10] : n10[10 - i];
  else if (i < 43) tab[i] = p10[i - 20];
  else if (i < kElfTab) tab[i] = n10[i - 43];
}

__host__ __device__ __forceinline__ double elf_recover_tab(double v, int32_t beta, const double *tab) {
  const double x = fabs(v);
  int32_t sp;
  if (x >= tab[0] && x < 1e9) {
    const int32_t e = (int32_t)((as_u64(x) >> 52) & 0x7ff) - 1023;
    const int32_t k0 = (e * 1233) >> 12;  // floor(e log10(2)); k0 + 1 in [-10, 9]
    sp = k0 + (x >= tab[k0 + 11] ? 1 : 0);
  } else {
    sp = elf_get_sp(x);
  }
  if (beta == 0) {
    const uint32_t k = 0u - (uint32_t)sp - 1u;
    double w = 0.0;
    if (k <= 22) w = tab[43 + k];
    else if (k < 400) w = c_n10[k];
    return v < 0.0 ? -w : w;
  }
  const uint32_t alpha = (uint32_t)beta - (uint32_t)sp - 1u;
  if (alpha <= 22 && x >= tab[0] && x < 1e9) {
    const double D = tab[20 + alpha], y = tab[43 + alpha];
    const double t = v * D;
    const double N = v < 0.0 ? floor(t) : ceil(t);
    return div_rn(N, D, y);
  }
  return elf_round_up(v, alpha);
}

// State of one XOR stream between calls of its decoder: the bit window and
// the values the next record depends on (Gorilla: v = previous value, a = L,
// b = T; Chimp: v, a = stored lead; Elf: v = stored bits, a = L, b = T,
// beta = last beta*).  A decoder called on records [i0, i1) leaves the
// state at record i1, so a segment can be decoded in several runs.
template <class W>
struct XorStateT {
  W r;
  uint64_t v;
  uint32_t a, b;
  int32_t beta;
  uint32_t end;  // set with kNanEnd: the record index of the NaN end marker
};
using XorState = XorStateT<BitWindow>;        // lockstep segments
using SerialState = XorStateT<SerialWindow>;  // lane-per-chunk runs (deep prefetch)

// Decoder state (all but the window) at a checkpoint written by the index
// pass (255 / -1 encode "unset").
template <class S>
__host__ __device__ __forceinline__ void xor_state_set(S &s, uint32_t codec, const XorCkpt &k) {
  s.v = k.value;
  if (codec == kCodecGorilla) {
    s.a = k.a;
    s.b = k.b;
  } else {
    s.a = k.a == 255 ? 0xffffffffu : k.a;
    s.b = k.b == 255 ? 0xffffffffu : k.b;
  }
  s.beta = k.beta < 0 ? INT32_MAX : (int32_t)k.beta;
}

// State at a checkpoint, reading the chunk's stream in place.
template <class S>
__host__ __device__ __forceinline__ void xor_state_at(S &s, uint32_t codec, const uint8_t *data, const XorCkpt &k) {
  s.r.init(data, k.bitpos);
  xor_state_set(s, codec, k);
}

// State before record 0 of a chunk.
template <class S>
__host__ __device__ __forceinline__ void xor_state_start(S &s, uint32_t codec, const uint8_t *data) {
  s.r.init(data, 0);
  s.v = 0;
  s.a = codec == kCodecGorilla ? 0u : 0xffffffffu;
  s.b = codec == kCodecGorilla ? 0u : 0xffffffffu;
  s.beta = INT32_MAX;
}

// Record sinks of the XOR decoders: put(i, v) receives record i of the chunk.
struct DirectSink {  // one lane per chunk: records stored as they come
  double *out;
  __host__ __device__ __forceinline__ void put(uint32_t i, double v) const { out[i] = v; }
};

// ------------------------------------------------------------------ Gorilla
// decoder/chunk_reader/gorilla.rs:77-157.  Decodes records [i0, i1) from
// state s (v = record i0-1, L, T); when i1 == n also checks that the next
// record is the NaN terminator.  ck (index pass): checkpoints every kSeg.
template <bool kWrite, class Sink, class S>
__host__ __device__ __forceinline__ uint32_t gorilla_run(S &s, uint32_t size_bits, uint32_t n, uint32_t i0,
                                                uint32_t i1, const Sink &out, XorCkpt *ck) {
  auto &r = s.r;
  uint64_t v = s.v;
  uint32_t L = s.a, T = s.b;
  const uint32_t last = i1 == n ? n + 1 : i1;
  uint32_t i = i0;
  if (i == 0 && i < last) {  // first value: 64 raw bits
    if (ck && n) ck[0] = XorCkpt{v, r.pos, (uint8_t)L, (uint8_t)T, -1, 0};
    v = r.field(0, 64);
    r.consume(64);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
    if (n == 0) return v == kNan ? kErrNone : kErrCount;
    if (v == kNan) {  // the chunk ends at record 0
      s.end = 0;
      return kNanEnd;
    }
    if (kWrite) out.put(0, as_f64(v));
    i = 1;
  }
  for (; i < last; ++i) {
    if (ck && i < n && (i % kSeg) == 0) ck[i / kSeg] = XorCkpt{v, r.pos, (uint8_t)L, (uint8_t)T, -1, 0};
    // control bit 0: repeat; 10: reuse (L, T); 11: 5-bit L, 6-bit sig - 1, then sig bits
    const uint32_t c2 = (uint32_t)(r.hi >> 62);
    const bool nz = c2 >= 2, fresh = c2 == 3;
    const uint32_t h = (uint32_t)(r.hi >> 51) & 0x7ff;
    const uint32_t l = fresh ? h >> 6 : L;
    const uint32_t sig = fresh ? (h & 63) + 1 : 64 - L - T;
    if (fresh && l + sig > 64) return kErrCorrupt;
    const uint32_t o = fresh ? 13 : 2;
    if (nz) {
      const uint32_t t = 64 - l - sig;
      v ^= r.field(o, sig) << t;
      L = l;
      T = t;
    }
    r.consume(nz ? o + sig : 1);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
    if (i == n) return v == kNan ? kErrNone : kErrCount;
    if (v == kNan) {  // the chunk ends at record i
      s.end = i;
      return kNanEnd;
    }
    if (kWrite) out.put(i, as_f64(v));
  }
  s.v = v;
  s.a = L;
  s.b = T;
  return kErrNone;
}

// -------------------------------------------------------------------- Chimp
// decoder/chunk_reader/chimp.rs:59-125.  State: v, stored lead (0xffffffff
// before the first 11/01 record).
template <bool kWrite, class Sink, class S>
__host__ __device__ __forceinline__ uint32_t chimp_run(S &s, uint32_t size_bits, uint32_t n, uint32_t i0,
                                              uint32_t i1, const Sink &out, XorCkpt *ck) {
  auto &r = s.r;
  uint64_t v = s.v;
  uint32_t lead = s.a;
  const uint32_t last = i1 == n ? n + 1 : i1;
  uint32_t i = i0;
  if (i == 0 && i < last) {  // first value: 64 raw bits
    if (ck && n) ck[0] = XorCkpt{v, r.pos, (uint8_t)(lead > 64 ? 255 : lead), 0, -1, 0};
    v = r.field(0, 64);
    r.consume(64);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
    if (n == 0) return v == kNan ? kErrNone : kErrCount;
    if (v == kNan) {  // the chunk ends at record 0
      s.end = 0;
      return kNanEnd;
    }
    if (kWrite) out.put(0, as_f64(v));
    i = 1;
  }
  for (; i < last; ++i) {
    if (ck && i < n && (i % kSeg) == 0)
      ck[i / kSeg] = XorCkpt{v, r.pos, (uint8_t)(lead > 64 ? 255 : lead), 0, -1, 0};
    // flag 00: same value; 11: 3-bit lead code, 64 - lead bits; 10: 64 - lead
    // bits with the stored lead; 01: 9-bit header (lead code, sig), sig bits
    const uint32_t f = (uint32_t)(r.hi >> 62);
    const uint32_t h9 = (uint32_t)(r.hi >> 53) & 0x1ff;
    const uint32_t c3 = h9 >> 6;  // bits 2..4: the lead code of flags 11 and 01
    const uint32_t nl = (f == 3 || f == 1) ? lead_repr(c3) : lead;
    if (f == 2 && lead > 64) return kErrCorrupt;
    uint32_t sig = 64 - nl;
    if (f == 1) {
      sig = (h9 & 63) ? (h9 & 63) : 64;
      if (sig + nl > 64) return kErrCorrupt;
    }
    const uint32_t o = f == 3 ? 5 : f == 1 ? 11 : 2;
    if (f) {
      v ^= r.field(o, sig) << (f == 1 ? 64 - sig - nl : 0);
      lead = nl;
    }
    r.consume(f ? o + sig : 2);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
    if (i == n) return v == kNan ? kErrNone : kErrCount;
    if (v == kNan) {  // the chunk ends at record i
      s.end = i;
      return kNanEnd;
    }
    if (kWrite) out.put(i, as_f64(v));
  }
  s.v = v;
  s.a = lead;
  return kErrNone;
}

// ---------------------------------------------------------------------- Elf
// ElfDecoderWrapper + ElfXorDecoder, decoder/chunk_reader/elf.rs:46-258.
// State: stored v' bits, L, T (0xffffffff before the first 10/11 record),
// last beta* (INT32_MAX before the first 11 prefix).
template <bool kWrite, class Sink, class S>
__host__ __device__ __forceinline__ uint32_t elf_run(S &s, uint32_t size_bits, uint32_t n, uint32_t i0,
                                            uint32_t i1, const Sink &out, XorCkpt *ck) {
  auto &r = s.r;
  uint64_t stored = s.v;
  uint32_t L = s.a, T = s.b;
  int32_t beta = s.beta;
  const uint32_t last = i1 == n ? n + 1 : i1;
  for (uint32_t i = i0; i < last; ++i) {
    if (ck && i < n && (i % kSeg) == 0)
      ck[i / kSeg] = XorCkpt{stored, r.pos, (uint8_t)(L > 64 ? 255 : L), (uint8_t)(T > 64 ? 255 : T),
                             (int8_t)(beta == INT32_MAX ? -1 : beta), 0};
    // prefix: 0 -> erased, reuse beta*; 10 -> not erased; 11 + 4 bits -> erased, new beta*
    const uint32_t p2 = (uint32_t)(r.hi >> 62);
    const uint32_t pre = (p2 & 2) ? (p2 == 2 ? 2u : 6u) : 1u;
    const bool recover = p2 != 2;
    if (p2 == 3) beta = (int32_t)((r.hi >> 58) & 15);
    bool term = false;
    uint32_t used;
    if (i == 0) {  // ElfXor first value: 7-bit trailing zeros, then 63 - tz bits
      const uint32_t tz = (uint32_t)r.field(pre, 7);
      if (tz < 64) {
        const uint32_t len = 63 - tz;
        const uint64_t m = len ? r.field(pre + 7, len) : 0;
        stored = ((m << 1) | 1) << tz;
        used = pre + 7 + len;
      } else {
        stored = 0;
        used = pre + 7;
      }
      term = stored == kNan;
    } else {
      // 11: 9-bit header (lead code, center), center - 1 bits; 10: 7-bit
      // header, center - 1 bits; 00: 64 - L - T bits with the stored L, T;
      // 01: same value
      const uint32_t f = (uint32_t)r.field(pre, 2);
      const uint32_t h9 = (uint32_t)r.field(pre + 2, 9), h7 = h9 >> 2;
      uint32_t nL = L, nT = T, len = 0, off = pre + 2;
      if (f == 3) {
        nL = lead_repr(h9 >> 6);
        const uint32_t c = (h9 & 63) ? (h9 & 63) : 64;
        if (nL + c > 64) return kErrCorrupt;
        nT = 64 - nL - c;
        len = c - 1;
        off = pre + 11;
      } else if (f == 2) {
        nL = lead_repr(h7 >> 4);
        const uint32_t c = (h7 & 15) ? (h7 & 15) : 16;
        if (nL + c > 64) return kErrCorrupt;
        nT = 64 - nL - c;
        len = c - 1;
        off = pre + 9;
      } else if (f == 0) {
        if ((uint64_t)L + T > 64) return kErrCorrupt;
        len = 64 - L - T;
      }
      const uint64_t m = len ? r.field(off, len) : 0;
      used = f == 1 ? pre + 2 : off + len;
      if (f != 1) {
        const uint64_t x = f == 0 ? m << T : ((m << 1) | 1) << nT;
        const uint64_t val = x ^ stored;
        if (val == kNan) term = true; else stored = val;
      }
      L = nL;
      T = nT;
    }
    r.consume(used);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
    if (i == n) return term ? kErrNone : kErrCount;
    if (term) {  // the chunk ends at record i
      s.end = i;
      return kNanEnd;
    }
    if (kWrite) {
      const double v = as_f64(stored);
      out.put(i, recover ? elf_recover(v, beta) : v);
    }
  }
  s.v = stored;
  s.a = L;
  s.b = T;
  s.beta = beta;
  return kErrNone;
}

// -------------------------------------------------------------- ElfOnChimp
// ElfDecoderWrapper over ChimpDecoder (decoder/chunk_reader/elf.rs:27-93
// with chimp.rs:59-125): per record the Elf prefix (0: erased, reuse beta*;
// 10: not erased; 11 + 4 bits: erased, new beta*), then one Chimp value
// (record 0: 64 raw bits; after it a 2-bit flag: 00 same, 01 lead code + sig
// + payload, 10 payload with the stored lead, 11 lead code + payload).  The
// chunk ends at the first Chimp value equal to NaN.  State: v = Chimp's
// stored value (the erased bits v'), a = stored lead (0xffffffff unset),
// beta = last beta* (INT32_MAX unset).
template <bool kWrite, class Sink, class S>
__host__ __device__ __forceinline__ uint32_t elfchimp_run(S &s, uint32_t size_bits, uint32_t n, uint32_t i0,
                                                 uint32_t i1, const Sink &out, XorCkpt *ck) {
  auto &r = s.r;
  uint64_t v = s.v;
  uint32_t lead = s.a;
  int32_t beta = s.beta;
  const uint32_t last = i1 == n ? n + 1 : i1;
  for (uint32_t i = i0; i < last; ++i) {
    if (ck && i < n && (i % kSeg) == 0)
      ck[i / kSeg] = XorCkpt{v, r.pos, (uint8_t)(lead > 64 ? 255 : lead), 0,
                             (int8_t)(beta == INT32_MAX ? -1 : beta), 0};
    const uint32_t p2 = (uint32_t)(r.hi >> 62);
    const uint32_t pre = (p2 & 2) ? (p2 == 2 ? 2u : 6u) : 1u;
    const bool recover = p2 != 2;
    if (p2 == 3) beta = (int32_t)((r.hi >> 58) & 15);
    uint32_t used;
    if (i == 0) {
      v = r.field(pre, 64);
      used = pre + 64;
    } else {
      const uint32_t f = (uint32_t)r.field(pre, 2);
      const uint32_t h9 = (uint32_t)r.field(pre + 2, 9);
      const uint32_t nl = (f == 3 || f == 1) ? lead_repr(h9 >> 6) : lead;
      if (f == 2 && lead > 64) return kErrCorrupt;
      uint32_t sig = 64 - nl;
      if (f == 1) {
        sig = (h9 & 63) ? (h9 & 63) : 64;
        if (sig + nl > 64) return kErrCorrupt;
      }
      const uint32_t o = pre + (f == 3 ? 5u : f == 1 ? 11u : 2u);
      if (f) {
        v ^= r.field(o, sig) << (f == 1 ? 64 - sig - nl : 0);
        lead = nl;
      }
      used = f ? o + sig : pre + 2;
    }
    r.consume(used);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
    if (i == n) return v == kNan ? kErrNone : kErrCount;
    if (v == kNan) {  // the chunk ends at record i
      s.end = i;
      return kNanEnd;
    }
    if (kWrite) {
      const double d = as_f64(v);
      out.put(i, recover ? elf_recover(d, beta) : d);
    }
  }
  s.v = v;
  s.a = lead;
  s.beta = beta;
  return kErrNone;
}

// Dispatch on the codec (wave-uniform).
template <bool kWrite, class Sink, class S>
__host__ __device__ __forceinline__ uint32_t xor_run(uint32_t codec, S &s, uint32_t size_bits, uint32_t n,
                                            uint32_t i0, uint32_t i1, const Sink &out, XorCkpt *ck) {
  switch (codec) {
    case kCodecGorilla: return gorilla_run<kWrite>(s, size_bits, n, i0, i1, out, ck);
    case kCodecChimp: return chimp_run<kWrite>(s, size_bits, n, i0, i1, out, ck);
    case kCodecElf: return elf_run<kWrite>(s, size_bits, n, i0, i1, out, ck);
    case kCodecElfOnChimp: return elfchimp_run<kWrite>(s, size_bits, n, i0, i1, out, ck);
    default: return kErrCorrupt;
  }
}

// ------------------------------------------------ lockstep segment runs
constexpr uint32_t kRun = 16;  // records of a run (between two flushes of the output stage)

// The segment kernel's form of the decoders above: the 64 lanes of a wave
// decode records a + k (k < cnt) of their own segments in one loop whose
// trip count cnt is the same on every lane, so the body is straight-line
// code (selects, no per-lane exits).  A lane takes part in record i while
// from <= i < lim and it has no error; lim = min(a + kRun, the segment's
// end), or n + 1 when the segment ends the chunk and its terminator n falls
// in [a, a + kRun] (the terminator is parsed and checked; cnt = kRun + 1
// when it is record a + kRun of some lane).  Record 0 of a chunk (64 raw
// bits) is decoded before, by xor_seg_first, which sets from = 1 on its
// lane.  A lane that meets an error keeps it,
// consumes no more bits and so reads no further than the forms above; the
// records it leaves in the sink are then meaningless.  sink.row(i, k, v,
// act) receives record i as row k of the run, act = whether the lane takes
// part in it (a sink may ignore the rows with act false).  Returns the
// lane's error.  kChecked false drops the per-record checks (corrupt
// header, overrun, terminator): for streams the load-time index has already
// parsed whole and accepted, where none can fire; the returned error is
// then the one passed in.

// Record 0 of a Gorilla or Chimp chunk (the stream's first 64 bits).
template <class Sink>
__host__ __device__ __forceinline__ uint32_t xor_seg_first(XorState &s, uint32_t size_bits, Sink &out) {
  BitWindow &r = s.r;
  s.v = r.field(0, 64);
  r.consume(64);
  r.refill();
  out.row(0, 0, as_f64(s.v), true);
  if (r.pos > size_bits) return kErrOverrun;
  return s.v == kNan ? kErrCount : kErrNone;  // n >= 1: record 0 is not the terminator
}

// Gorilla, decoder/chunk_reader/gorilla.rs:77-157 (the record step of
// gorilla_run).
template <bool kChecked, class Sink>
__host__ __device__ __forceinline__ uint32_t gorilla_seg(XorState &s, uint32_t size_bits, uint32_t n, uint32_t a,
                                                uint32_t cnt, uint32_t from, uint32_t lim, Sink &out,
                                                uint32_t e) {
  BitWindow &r = s.r;
  uint64_t v = s.v;
  uint32_t L = s.a, T = s.b;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t i = a + k;
    const bool act = (!kChecked || e == 0) && i >= from && i < lim;
    const uint32_t c2 = (uint32_t)(r.hi >> 62);
    const bool nz = c2 >= 2, fresh = c2 == 3;
    const uint32_t h = (uint32_t)(r.hi >> 51) & 0x7ff;
    const uint32_t l = fresh ? h >> 6 : L;
    const uint32_t sig = fresh ? (h & 63) + 1 : 64 - L - T;
    const bool bad = kChecked && fresh && l + sig > 64;
    const uint32_t o = fresh ? 13 : 2;
    const uint32_t t = (64 - l - sig) & 63;
    const uint64_t x = r.field(o, sig) << t;
    const bool upd = act && nz && !bad;
    v ^= upd ? x : 0;
    L = upd ? l : L;
    T = upd ? t : T;
    r.consume(act && !bad ? (nz ? o + sig : 1u) : 0u);
    r.refill();
    if (kChecked) {
      const uint32_t code = bad                      ? (uint32_t)kErrCorrupt
                            : r.pos > size_bits      ? (uint32_t)kErrOverrun
                            : (i == n) != (v == kNan) ? (uint32_t)kErrCount
                                                     : 0u;
      e = act ? code : e;
    }
    out.row(i, k, as_f64(v), act);
  }
  s.v = v;
  s.a = L;
  s.b = T;
  return e;
}

// Chimp, decoder/chunk_reader/chimp.rs:59-125 (the record step of chimp_run).
template <bool kChecked, class Sink>
__host__ __device__ __forceinline__ uint32_t chimp_seg(XorState &s, uint32_t size_bits, uint32_t n, uint32_t a,
                                              uint32_t cnt, uint32_t from, uint32_t lim, Sink &out,
                                              uint32_t e) {
  BitWindow &r = s.r;
  uint64_t v = s.v;
  uint32_t lead = s.a;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t i = a + k;
    const bool act = (!kChecked || e == 0) && i >= from && i < lim;
    const uint32_t f = (uint32_t)(r.hi >> 62);
    const uint32_t h9 = (uint32_t)(r.hi >> 53) & 0x1ff;
    const uint32_t nl = (f & 1) ? lead_repr(h9 >> 6) : lead;  // flags 11 and 01 carry a lead code
    const uint32_t s6 = h9 & 63;
    const uint32_t sig = f == 1 ? (s6 ? s6 : 64u) : 64 - nl;
    const bool bad = kChecked && ((f == 2 && lead > 64) || (f == 1 && sig + nl > 64));
    const uint32_t o = f == 3 ? 5u : f == 1 ? 11u : 2u;
    const uint64_t x = r.field(o, sig) << (f == 1 ? (64 - sig - nl) & 63 : 0u);
    const bool upd = act && f != 0 && !bad;
    v ^= upd ? x : 0;
    lead = upd ? nl : lead;
    r.consume(act && !bad ? (f ? o + sig : 2u) : 0u);
    r.refill();
    if (kChecked) {
      const uint32_t code = bad                      ? (uint32_t)kErrCorrupt
                            : r.pos > size_bits      ? (uint32_t)kErrOverrun
                            : (i == n) != (v == kNan) ? (uint32_t)kErrCount
                                                     : 0u;
      e = act ? code : e;
    }
    out.row(i, k, as_f64(v), act);
  }
  s.v = v;
  s.a = lead;
  return e;
}

// Record 0 of an Elf chunk: prefix, then the ElfXor first value (7-bit
// trailing zeros, 63 - tz bits), decoder/chunk_reader/elf.rs:105-140.
template <class Sink>
__host__ __device__ __forceinline__ uint32_t elf_seg_first(XorState &s, uint32_t size_bits, Sink &out,
                                                  const double *tab) {
  BitWindow &r = s.r;
  const uint32_t p2 = (uint32_t)(r.hi >> 62);
  const uint32_t pre = (p2 & 2) ? (p2 == 2 ? 2u : 6u) : 1u;
  if (p2 == 3) s.beta = (int32_t)((r.hi >> 58) & 15);
  const uint32_t tz = (uint32_t)r.field(pre, 7);
  uint32_t used = pre + 7;
  s.v = 0;
  if (tz < 64) {
    const uint32_t len = 63 - tz;
    const uint64_t m = len ? r.field(pre + 7, len) : 0;
    s.v = ((m << 1) | 1) << tz;
    used += len;
  }
  r.consume(used);
  r.refill();
  const double v = as_f64(s.v);
  out.row(0, 0, p2 != 2 ? elf_recover_tab(v, s.beta, tab) : v, true);
  if (r.pos > size_bits) return kErrOverrun;
  return s.v == kNan ? kErrCount : kErrNone;  // n >= 1: record 0 is not the terminator
}

// Elf, decoder/chunk_reader/elf.rs:46-258 (the record step of elf_run for
// records i >= 1).  Only records the lane takes part in reach the sink.
// tab: the powers of ten of elf_recover_tab.
template <bool kChecked, class Sink>
__host__ __device__ __forceinline__ uint32_t elf_seg(XorState &s, uint32_t size_bits, uint32_t n, uint32_t a,
                                            uint32_t cnt, uint32_t from, uint32_t lim, Sink &out,
                                            uint32_t e, const double *tab) {
  BitWindow &r = s.r;
  uint64_t stored = s.v;
  uint32_t L = s.a, T = s.b;
  int32_t beta = s.beta;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t i = a + k;
    const bool act = (!kChecked || e == 0) && i >= from && i < lim;
    // prefix (0: erased, reuse beta*; 10: not erased; 11 + 4 bits: erased,
    // new beta*), then the 2-bit XOR flag and a 9-bit header, all within
    // the window's first 17 bits
    const uint32_t h32 = (uint32_t)(r.hi >> 32);
    const uint32_t p2 = h32 >> 30;
    const uint32_t pre = p2 == 3 ? 6u : p2 == 2 ? 2u : 1u;
    const int32_t nbeta = p2 == 3 ? (int32_t)((h32 >> 26) & 15) : beta;
    const uint32_t f = (h32 >> (30 - pre)) & 3;
    const uint32_t h9 = (h32 >> (21 - pre)) & 0x1ff;
    // 11: 3-bit lead code + 6-bit center; 10: 3-bit lead code + 4-bit
    // center; 00: the stored L, T; 01: same value.  Plain selects (the
    // compiler keeps them branch-free).
    const bool f3 = f == 3, hdr = f >= 2, f0 = f == 0;
    const uint32_t nL = hdr ? lead_repr(h9 >> 6) : L;
    const uint32_t cb = f3 ? h9 & 63 : (h9 >> 2) & 15;
    const uint32_t c = cb ? cb : (f3 ? 64u : 16u);
    const bool bad = kChecked && ((hdr & (nL + c > 64)) | (f0 & ((uint64_t)L + T > 64)));
    const uint32_t nT = hdr ? (64 - nL - c) & 63 : T;
    const uint32_t len = hdr ? c - 1 : (f0 ? 64 - L - T : 0u);
    const uint32_t off = pre + (f3 ? 11u : hdr ? 9u : 2u);
    const uint64_t m = r.field(off, len ? len : 1u) & (len ? ~0ull : 0ull);
    const uint64_t x = f == 0 ? m << (T & 63) : ((m << 1) | 1) << nT;
    const uint64_t val = x ^ stored;
    const bool term = kChecked && f != 1 && val == kNan;  // unchecked: only record n can be the NaN, and it is never output
    const bool upd = act && !bad;
    stored = upd && f != 1 && !term ? val : stored;
    L = upd ? nL : L;
    T = upd ? nT : T;
    beta = upd ? nbeta : beta;
    r.consume(upd ? (f == 1 ? pre + 2 : off + len) : 0u);
    r.refill();
    if (kChecked) {
      const uint32_t code = bad                 ? (uint32_t)kErrCorrupt
                            : r.pos > size_bits ? (uint32_t)kErrOverrun
                            : (i == n) != term  ? (uint32_t)kErrCount
                                                : 0u;
      e = act ? code : e;
    }
    const double v = as_f64(stored);
    const double w = p2 != 2 ? elf_recover_tab(v, beta, tab) : v;
    out.row(i, k, w, act);
  }
  s.v = stored;
  s.a = L;
  s.b = T;
  s.beta = beta;
  return e;
}

// Record 0 of an ElfOnChimp chunk: prefix, then Chimp's 64 raw bits.
template <class Sink>
__host__ __device__ __forceinline__ uint32_t elfchimp_seg_first(XorState &s, uint32_t size_bits, Sink &out,
                                                       const double *tab) {
  BitWindow &r = s.r;
  const uint32_t p2 = (uint32_t)(r.hi >> 62);
  const uint32_t pre = (p2 & 2) ? (p2 == 2 ? 2u : 6u) : 1u;
  if (p2 == 3) s.beta = (int32_t)((r.hi >> 58) & 15);
  s.v = r.field(pre, 64);
  r.consume(pre + 64);
  r.refill();
  const double v = as_f64(s.v);
  out.row(0, 0, p2 != 2 ? elf_recover_tab(v, s.beta, tab) : v, true);
  if (r.pos > size_bits) return kErrOverrun;
  return s.v == kNan ? kErrCount : kErrNone;  // n >= 1: record 0 is not the end marker
}

// ElfOnChimp, the record step of elfchimp_run for records i >= 1 in the
// lockstep form: the Elf prefix, then chimp_seg's step at offset `pre`
// (the 2-bit flag and the 9-bit header lie in the window's first 17 bits).
template <bool kChecked, class Sink>
__host__ __device__ __forceinline__ uint32_t elfchimp_seg(XorState &s, uint32_t size_bits, uint32_t n, uint32_t a,
                                                 uint32_t cnt, uint32_t from, uint32_t lim, Sink &out,
                                                 uint32_t e, const double *tab) {
  BitWindow &r = s.r;
  uint64_t v = s.v;
  uint32_t lead = s.a;
  int32_t beta = s.beta;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t i = a + k;
    const bool act = (!kChecked || e == 0) && i >= from && i < lim;
    const uint32_t h32 = (uint32_t)(r.hi >> 32);
    const uint32_t p2 = h32 >> 30;
    const uint32_t pre = p2 == 3 ? 6u : p2 == 2 ? 2u : 1u;
    const int32_t nbeta = p2 == 3 ? (int32_t)((h32 >> 26) & 15) : beta;
    const uint32_t f = (h32 >> (30 - pre)) & 3;
    const uint32_t h9 = (h32 >> (21 - pre)) & 0x1ff;
    const uint32_t nl = (f & 1) ? lead_repr(h9 >> 6) : lead;  // flags 11 and 01 carry a lead code
    const uint32_t s6 = h9 & 63;
    const uint32_t sig = f == 1 ? (s6 ? s6 : 64u) : 64 - nl;
    const bool bad = kChecked && ((f == 2 && lead > 64) || (f == 1 && sig + nl > 64));
    const uint32_t o = pre + (f == 3 ? 5u : f == 1 ? 11u : 2u);
    const uint64_t x = r.field(o, sig) << (f == 1 ? (64 - sig - nl) & 63 : 0u);
    const bool upd = act && !bad;
    v ^= upd && f != 0 ? x : 0;
    lead = upd && f != 0 ? nl : lead;
    beta = upd ? nbeta : beta;
    r.consume(upd ? (f ? o + sig : pre + 2) : 0u);
    r.refill();
    if (kChecked) {
      const uint32_t code = bad                      ? (uint32_t)kErrCorrupt
                            : r.pos > size_bits      ? (uint32_t)kErrOverrun
                            : (i == n) != (v == kNan) ? (uint32_t)kErrCount
                                                     : 0u;
      e = act ? code : e;
    }
    const double d = as_f64(v);
    out.row(i, k, p2 != 2 ? elf_recover_tab(d, beta, tab) : d, act);
  }
  s.v = v;
  s.a = lead;
  s.beta = beta;
  return e;
}

// Chimp128 parse, decoder/chunk_reader/chimp_n.rs:173-285, in the lockstep
// form, over one segment of kSeg128 records per lane.  Record i is parsed
// as (source distance, XOR payload): value_i = value_{i - dist} ^ x.  Flags
// 10 / 11 use the previous record (dist 1); flags 00 / 01 name ring slot
// idx, i.e. the latest record j <= i - 1 with j == idx (mod 128), so dist =
// i - j is in [1, 128].  dist 0: no source -- record 0 (x = its 64 raw bits)
// or an unwritten ring slot (value 0) -- so value_i = x.  No values are
// formed here (that needs the ring): the segment kernel resolves them after
// the parse.  Within the lane's segment a record with dist 1 is folded into
// its predecessor (value_i = value_{i-1} ^ x_i = value_{src(i-1)} ^
// (x'_{i-1} ^ x_i)): it is passed on as x' = x'_{i-1} ^ x_i with the
// predecessor's source (dist' = dist'_{i-1} + 1, or 0 when the predecessor
// has none; at most 128 + 15), so the resolve follows a few
// segment-to-segment links instead of record chains.  Record i (i <= n, the
// terminator n included) goes to the sink: out.row(i, k, x', act) and
// out.dist(k, dist', act) (k = i - a).  Record 0 is parsed before, by
// chimp128_seg_first, which starts the fold state (px, pd).
template <class Sink>
__host__ __device__ __forceinline__ uint32_t chimp128_seg_first(BitWindow &r, uint32_t size_bits, Sink &out,
                                                       uint64_t &px, uint32_t &pd) {
  const uint64_t x = r.field(0, 64);
  r.consume(64);
  r.refill();
  out.row(0, 0, as_f64(x), true);
  out.dist(0, 0, true);
  px = x;
  pd = 0;
  return r.pos > size_bits ? (uint32_t)kErrOverrun : 0u;
}

// px / pd: x' and dist' of the record before `a` when it is in this segment
// (have_prev), then of each record parsed.
template <bool kChecked, class Sink>
__host__ __device__ __forceinline__ uint32_t chimp128_parse_seg(BitWindow &r, uint32_t &lead, uint32_t size_bits,
                                                       uint32_t a, uint32_t cnt, uint32_t from, uint32_t lim,
                                                       Sink &out, uint32_t e, uint64_t px, uint32_t pd,
                                                       bool have_prev) {
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t i = a + k;
    const bool act = (!kChecked || e == 0) && i >= from && i < lim;
    // 00: 7-bit ring index; 11: lead code, 64 - lead bits; 10: 64 - lead
    // bits with the stored lead; 01: 16-bit header (index, lead code, sig)
    const uint32_t f = (uint32_t)(r.hi >> 62);
    const uint32_t h = (uint32_t)(r.hi >> 46) & 0xffff;  // bits 2..17
    const uint32_t idx = h >> 9;
    const uint32_t lc = lead_repr(f == 3 ? h >> 13 : (h >> 6) & 7);  // flags 11 and 01 carry a lead code
    const uint32_t nl = (f & 1) ? lc : lead;
    const uint32_t s6 = h & 63;
    const uint32_t len = f == 1 ? (s6 ? s6 : 64u) : 64 - nl;
    const bool bad = kChecked && (((f == 2) & (lead > 64)) | ((f == 1) & (len + nl > 64)));
    const uint32_t off = f == 3 ? 5u : f == 1 ? 18u : 2u;
    const uint64_t x = r.field(off, len) << (f == 1 ? (64 - len - nl) & 63 : 0u) & (f ? ~0ull : 0ull);
    const uint32_t back = f <= 1 ? (((i - 1) - idx) & 127u) + 1 : 1u;  // i - source
    const uint32_t dist0 = back > i ? 0u : back;
    const bool fold = have_prev && dist0 == 1;  // the source is this segment's previous record
    const uint64_t xf = fold ? x ^ px : x;
    const uint32_t dist = fold ? (pd ? pd + 1 : 0u) : dist0;
    px = act ? xf : px;
    pd = act ? dist : pd;
    have_prev = have_prev || act;
    const bool upd = act && !bad;
    lead = upd && f != 0 ? nl : lead;
    r.consume(upd ? (f == 0 ? 9u : off + len) : 0u);
    r.refill();
    if (kChecked) {
      const uint32_t code = bad ? (uint32_t)kErrCorrupt : r.pos > size_bits ? (uint32_t)kErrOverrun : 0u;
      e = act ? code : e;
    }
    out.row(i, k, as_f64(xf), act);
    out.dist(k, dist, act);
  }
  return e;
}

// Output staging of a wave decoding 64 checkpoint segments (segment
// first + l on lane l) in runs of kRun = 16 records: record i of lane l
// goes to slot (i mod 16, l ^ stage_swz(i mod 16)) of the wave's [16][64]
// array; the XOR swizzle keeps both the per-record writes (one row, all
// lanes) and the flush reads (pairs of rows, 8 lanes) free of LDS bank
// conflicts beyond the two passes a 64-bit access takes.
__device__ __forceinline__ uint32_t stage_swz(uint32_t r) { return (r >> 1) << 2; }

struct StageSink {
  double *stage;
  uint8_t *dists;  // Chimp128 jobs: the source distances, job record t at dists[t]
  __device__ __forceinline__ void put(uint32_t i, double v) const {
    const uint32_t r = i & (kRun - 1);
    stage[r * kWave + (lane_id() ^ stage_swz(r))] = v;
  }
  // lockstep runs: record i is row k of the run (k <= kRun; row kRun holds
  // only a terminator, never flushed); rows the lane takes no part in keep
  // what they held
  __device__ __forceinline__ void row(uint32_t, uint32_t k, double v, bool act = true) const {
    if (act) stage[k * kWave + (lane_id() ^ stage_swz(k))] = v;
  }
  // a Chimp128 lane's record k: job record lane * kSeg128 + k
  __device__ __forceinline__ void dist(uint32_t k, uint32_t d, bool act) const {
    if (act) dists[lane_id() * kSeg128 + k] = (uint8_t)d;
  }
};

// Writes the staged records k0 .. k0 + 15 of the wave's 64 segments to dst
// (record 0 of segment `first`), called by every lane of the wave: store k
// covers segments 8k .. 8k + 7, lane j the records k0 + 2(j % 8) and
// k0 + 2(j % 8) + 1 of segment 8k + j / 8, so each store instruction writes
// eight whole 128-byte lines.  Records at or past n_job (the chunk's end)
// are not stored.
__device__ __forceinline__ void stage_flush(const double *stage, double *dst, uint32_t k0, uint32_t n_job) {
  __builtin_amdgcn_wave_barrier();
  const uint32_t j = lane_id();
  const uint32_t r = 2 * (j & 7);
  const bool aligned = (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; ++k) {
    const uint32_t p = k * 8 + (j >> 3);
    const uint32_t g = p * kSeg + k0 + r;
    const uint32_t x = p ^ stage_swz(r);
    const double a = stage[r * kWave + x], b = stage[(r + 1) * kWave + x];
    if (g + 1 < n_job) {
      if (aligned) {
        st_pair(dst + g, a, b);
      } else {
        st_one(dst + g, a);
        st_one(dst + g + 1, b);
      }
    } else if (g < n_job) {
      st_one(dst + g, a);
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// ----------------------------------------------------------------- Chimp128
// decoder/chunk_reader/chimp_n.rs:173-285.  `ring` is this lane's 128-slot
// column in LDS, stride kWave (slot s of lane l at ring[s * 64]).
// A NaN before record n ends the chunk there: kNanEnd, its index in *end.
template <bool kWrite>
__device__ uint32_t chimp128_run(const uint8_t *data, uint32_t size_bits, uint32_t n, double *out,
                                 uint64_t *ring, uint32_t *end) {
  StreamReader r;
  r.init(data, 0);
  for (int s = 0; s < 128; ++s) ring[s * kWave] = 0;
  uint32_t cur = 0, lead = 0xffffffffu;
  uint64_t v = 0;
  for (uint32_t i = 0; i <= n; ++i) {
    bool term = false;
    if (i == 0) {
      v = r.read64(64);
      ring[0] = v;
      term = v == kNan;
    } else {
      const uint32_t f = r.read(2);
      if (f == 0) {
        v = ring[r.read(7) * kWave];
        cur = (cur + 1) & 127;
        ring[cur * kWave] = v;
      } else {
        if (f == 3) {
          lead = lead_repr(r.read(3));
          v ^= r.read64(64 - lead);
        } else if (f == 2) {
          if (lead > 64) return kErrCorrupt;
          v ^= r.read64(64 - lead);
        } else {
          const uint32_t h = r.read(16);
          const uint32_t idx = (h >> 9) & 127;
          lead = lead_repr((h >> 6) & 7);
          uint32_t sig = h & 63;
          if (!sig) sig = 64;
          if (sig + lead > 64) return kErrCorrupt;
          v = ring[idx * kWave] ^ (r.read64(sig) << (64 - sig - lead));
        }
        if (v == kNan) {
          term = true;
        } else {
          cur = (cur + 1) & 127;
          ring[cur * kWave] = v;
        }
      }
    }
    if (r.pos > size_bits) return kErrOverrun;
    if (i == n) return term ? kErrNone : kErrCount;
    if (term) {
      *end = i;
      return kNanEnd;
    }
    if (kWrite) out[i] = as_f64(v);
  }
  return kErrNone;
}

// Chimp128 parse-only pass for the load-time index: record lengths depend
// only on flags, lead codes and the stored lead, so one lane walks records
// 0..n (n = the terminator) without the value ring and writes a checkpoint
// {bit position, stored lead} every kSeg128 records (n / kSeg128 + 1
// entries).
// One Chimp128 record (i >= 1) at the start of window r
// (decoder/chunk_reader/chimp_n.rs:189-285): flag 00 = 7-bit ring index;
// 11 = 3-bit lead code, 64 - lead bits; 10 = 64 - lead bits with the stored
// lead; 01 = 16-bit header (index, lead code, sig), sig bits.  Sets the
// flag, ring index, XOR payload (aligned) and bits used; false on a header
// the reference would reject.
template <class W>
__host__ __device__ __forceinline__ bool chimp128_rec(const W &r, uint32_t &lead, uint32_t &f, uint32_t &idx,
                                             uint64_t &x, uint32_t &used) {
  f = (uint32_t)(r.hi >> 62);
  const uint32_t h = (uint32_t)(r.hi >> 46) & 0xffff;  // bits 2..17
  idx = h >> 9;
  if (f == 0) {
    x = 0;
    used = 9;
    return true;
  }
  uint32_t nl = lead, len, off;
  if (f == 3) {
    nl = lead_repr(h >> 13);
    len = 64 - nl;
    off = 5;
  } else if (f == 2) {
    if (lead > 64) return false;
    len = 64 - nl;
    off = 2;
  } else {
    idx &= 127;
    nl = lead_repr((h >> 6) & 7);
    len = (h & 63) ? (h & 63) : 64;
    if (len + nl > 64) return false;
    off = 18;
  }
  x = r.field(off, len) << (f == 1 ? 64 - len - nl : 0);
  lead = nl;
  used = off + len;
  return true;
}

__host__ __device__ inline uint32_t chimp128_index(const uint8_t *data, uint32_t size_bits, uint32_t n, XorCkpt *ck) {
  SerialWindow r;
  r.init(data, 0);
  uint32_t lead = 0xffffffffu;
  for (uint32_t i = 0; i <= n; ++i) {
    if ((i % kSeg128) == 0) ck[i / kSeg128] = XorCkpt{0, r.pos, (uint8_t)(lead > 64 ? 255 : lead), 0, -1, 0};
    uint32_t used = 64;
    if (i) {
      uint32_t f, idx;
      uint64_t x;
      if (!chimp128_rec(r, lead, f, idx, x, used)) return kErrCorrupt;
    }
    r.consume(used);
    r.refill();
    if (r.pos > size_bits) return kErrOverrun;
  }
  return kErrNone;
}
// -------------------------------------------------------------- tile codecs
// Records of tile t handled by this lane: [t*kTile + lane*8, +8) clipped to n.

// Uncompressed, chunk_reader/uncompressed.rs:56-71: a copy; lanes move
// 16-byte pairs so each wave instruction covers 1 KiB contiguous.
__device__ __forceinline__ void plain_tile(const DevChunk &c, uint32_t t, double *out) {
  const EBI_GLOBAL double *src = as_global(reinterpret_cast<const double *>(c.data));
  const uint32_t lane = lane_id();
  const uint32_t b = t * kTile;
  const bool pair_ok = ((c.rec_off + (uint64_t)reinterpret_cast<uintptr_t>(out) / 8) & 1) == 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = b + k * 128 + lane * 2;
    if (pair_ok && i + 1 < c.n) {
      const double2 x = *reinterpret_cast<const EBI_GLOBAL double2 *>(src + i);
      st_pair(out + c.rec_off + i, x.x, x.y);
    } else {
      if (i < c.n) st_one(out + c.rec_off + i, src[i]);
      if (i + 1 < c.n) st_one(out + c.rec_off + i + 1, src[i + 1]);
    }
  }
}

// BUFF deltas of this lane's 8 records: planes MSB-first (each byte XOR 0x80),
// then the LSB-first tail of r bits (compression_common/buff.rs:85-287,
// chunk_reader/buff/decode.rs:5-70).  Lanes whose records lie past n get
// garbage-free zeros.  Controlled precision reads only the first np planes
// and the tail if `tail` (the bits not read are 0).
__device__ __forceinline__ void buff_deltas(const DevChunk &c, uint32_t i0, uint64_t d[8], uint32_t np, bool tail) {
#pragma unroll
  for (int j = 0; j < 8; ++j) d[j] = 0;
  if (i0 >= c.n) return;
  const EBI_GLOBAL uint8_t *base = as_global(reinterpret_cast<const uint8_t *>(c.data));
  for (uint32_t p = 0; p < np; ++p) {
    const uint64_t w = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(base + (size_t)p * c.plane_stride + i0);
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] = (d[j] << 8) | (((w >> (8 * j)) & 0xff) ^ 0x80);
  }
  if (np < c.n_planes) {  // np >= 1 here, so the shift is below 64
    const uint32_t sh = 8 * (c.n_planes - np) + c.r;
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] <<= sh;
    return;
  }
  if (c.r && !tail) {
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] <<= c.r;
    return;
  }
  if (c.r) {
    // 8 records x r bits = r bytes starting at byte i0*r/8 (byte aligned).
    const EBI_GLOBAL uint8_t *tail = base + (size_t)c.n_planes * c.plane_stride;
    const uint32_t byte0 = (i0 / 8) * c.r;
    const uint32_t a = byte0 & ~7u, sh = (byte0 & 7u) * 8;
    const uint64_t lo = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(tail + a);
    const uint64_t hi = *reinterpret_cast<const EBI_GLOBAL uint64_t *>(tail + a + 8);
    const uint64_t bits = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
    const uint64_t m = (1ull << c.r) - 1;
#pragma unroll
    for (int j = 0; j < 8; ++j) d[j] = (d[j] << c.r) | ((bits >> (j * c.r)) & m);
  }
}

__device__ __forceinline__ void store8(double *dst, const double v[8], uint32_t cnt, bool aligned) {
  if (aligned && cnt == 8) {
    double2 *p = reinterpret_cast<double2 *>(dst);
    p[0] = make_double2(v[0], v[1]);
    p[1] = make_double2(v[2], v[3]);
    p[2] = make_double2(v[4], v[5]);
    p[3] = make_double2(v[6], v[7]);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if ((uint32_t)j < cnt) dst[j] = v[j];
  }
}

// (base + delta) as f64 / 2^frac; ldexp is exact like the reference's divide.
__device__ __forceinline__ double buff_value(const DevChunk &c, uint64_t d) {
  return ldexp((double)(int64_t)((uint64_t)c.base + d), -(int)c.frac);
}

// Stores a tile whose lane l holds records 8l .. 8l+7 (v[]) as coalesced
// 16-byte stores: the wave transposes through `stage` (kStageStride doubles
// per lane), then store k of lane l covers records 128k + 2l and 128k + 2l + 1,
// so each store instruction writes 1 KiB contiguous.  The 80-byte lane stride
// keeps both the 16-byte LDS writes and reads free of bank conflicts.  Every
// lane of the wave must call it (it synchronises the wave).  nv = records of
// the tile to write (<= kTile).
constexpr uint32_t kStageStride = 10;
__device__ __forceinline__ void tile_store(double *dst, const double v[8], uint32_t nv, double *stage) {
  const uint32_t lane = lane_id();
  double2 *mine = reinterpret_cast<double2 *>(stage + lane * kStageStride);
  mine[0] = make_double2(v[0], v[1]);
  mine[1] = make_double2(v[2], v[3]);
  mine[2] = make_double2(v[4], v[5]);
  mine[3] = make_double2(v[6], v[7]);
  wave_sync();
  const bool aligned = (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint32_t r = 128 * k + 2 * lane;
    const double2 x = *reinterpret_cast<const double2 *>(stage + (r >> 3) * kStageStride + (r & 7));
    if (aligned && r + 1 < nv) {
      st_pair(dst + r, x.x, x.y);
    } else {
      if (r < nv) st_one(dst + r, x.x);
      if (r + 1 < nv) st_one(dst + r + 1, x.y);
    }
  }
  wave_sync();  // the stage is reused by the wave's next job
}

__device__ __forceinline__ void buff_deltas(const DevChunk &c, uint32_t i0, uint64_t d[8]) {
  buff_deltas(c, i0, d, c.n_planes, true);
}

// PRECISION_MAP (compression_common/buff.rs:36): fractional bits kept for
// 0..12 decimal digits, six bits each.
__host__ __device__ __forceinline__ uint32_t buff_precision_bits(uint32_t p) {
  constexpr uint64_t lo = 0ull | (5ull << 6) | (8ull << 12) | (11ull << 18) | (15ull << 24) | (18ull << 30) |
                          (21ull << 36) | (25ull << 42) | (28ull << 48) | (31ull << 54);
  constexpr uint64_t hi = 35ull | (38ull << 6) | (50ull << 12);
  return (uint32_t)((p < 10 ? lo >> (6 * p) : hi >> (6 * (p - 10))) & 63);
}

// Planes and tail that decode_with_precision reads
// (chunk_reader/buff/controlled_precision.rs:21-75): with U = frac minus the
// bits kept for `precision` (0 if fewer), plane 0 always, plane p >= 1 while
// the B - 8p bits below it exceed U, the tail (r bits) only after every
// plane and if r > U.  B < 8 decodes every bit.  precision < 0: full.
__host__ __device__ __forceinline__ void buff_precision_reads(const DevChunk &c, int32_t precision, uint32_t *np,
                                                              bool *tail) {
  *np = c.n_planes;
  *tail = true;
  if (precision < 0 || c.bits < 8) return;
  const uint32_t keep = buff_precision_bits((uint32_t)precision);
  const uint32_t u = c.frac > keep ? c.frac - keep : 0u;
  uint32_t p = 1;
  while (p < c.n_planes && c.bits - 8 * p > u) ++p;
  *np = p;
  *tail = p == c.n_planes && c.r > u;
}

__device__ __forceinline__ void buff_tile(const DevChunk &c, uint32_t t, double *out, double *stage,
                                          int32_t precision) {
  const uint32_t i0 = t * kTile + lane_id() * 8;
  uint32_t np;
  bool tail;
  buff_precision_reads(c, precision, &np, &tail);
  uint64_t d[8];
  buff_deltas(c, i0, d, np, tail);  // zeros past the chunk's end
  double v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = buff_value(c, d[j]);
  tile_store(out + c.rec_off + (uint64_t)t * kTile, v, min(kTile, c.n - t * kTile), stage);
}

// DeltaSprintz zigzag deltas of this lane's 8 records (MSB-first W-bit
// fields, chunk_reader/sprintz.rs:480-504), unzigzagged; 0 past n.
__device__ __forceinline__ void sprintz_deltas(const DevChunk &c, uint32_t i0, int64_t d[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) d[j] = 0;
  if (i0 >= c.n || c.bits == 0) return;
  const uint32_t W = c.bits;
  const uint32_t cnt = min(8u, c.n - i0);
  if (W <= 16) {
    // 8 records x W bits = W bytes from byte i0 * W / 8: two aligned 16-byte
    // loads, a byte-aligned 128-bit funnel shift, then fixed-position fields
    const uint32_t start = (i0 / 8) * W;
    const EBI_GLOBAL uint4 *q = as_global(reinterpret_cast<const uint4 *>(c.data)) + (start >> 4);
    const uint4 a = q[0], b = q[1];
    const uint32_t sb = start & 15, wi = sb >> 2, sh = (sb & 3) * 8;
    // big-endian word wi + k of the 32 loaded bytes
    auto word = [&](uint32_t k) {
      const uint32_t i = wi + k;
      const uint32_t x = i == 0 ? a.x : i == 1 ? a.y : i == 2 ? a.z : i == 3 ? a.w
                       : i == 4 ? b.x : i == 5 ? b.y : i == 6 ? b.z : b.w;
      return __builtin_bswap32(x);
    };
    uint32_t u0 = word(0), u1 = word(1), u2 = word(2), u3 = word(3), u4 = word(4);
    if (sh) {
      u0 = (u0 << sh) | (u1 >> (32 - sh));
      u1 = (u1 << sh) | (u2 >> (32 - sh));
      u2 = (u2 << sh) | (u3 >> (32 - sh));
      u3 = (u3 << sh) | (u4 >> (32 - sh));
    }
    uint64_t hi = ((uint64_t)u0 << 32) | u1, lo = ((uint64_t)u2 << 32) | u3;
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // field j = the top W bits, then shift the pair by W
      const uint64_t z = hi >> (64 - W);
      if ((uint32_t)j < cnt) d[j] = (int64_t)(z >> 1) ^ -(int64_t)(z & 1);
      hi = (hi << W) | (lo >> (64 - W));
      lo <<= W;
    }
    return;
  }
  LaneReader r;
  r.init(reinterpret_cast<const uint8_t *>(c.data), i0 * W);
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if ((uint32_t)j < cnt) {
      const uint64_t z = r.read64(W);
      d[j] = (int64_t)(z >> 1) ^ -(int64_t)(z & 1);
    }
  }
}

// Inclusive quantized values q_i of this lane's records given the tile's
// starting prefix q_{t*kTile - 1} (or q0 - 0 for tile 0; the first delta is 0).
__device__ __forceinline__ void sprintz_q(const DevChunk &c, uint32_t t, int64_t tile_base, int64_t q[8]) {
  const uint32_t i0 = t * kTile + lane_id() * 8;
  int64_t d[8];
  sprintz_deltas(c, i0, d);
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    s = (int64_t)((uint64_t)s + (uint64_t)d[j]);
    q[j] = s;
  }
  const int64_t incl = wave_incl_scan_i64(s);
  const int64_t excl = (int64_t)((uint64_t)incl - (uint64_t)s) + tile_base;
#pragma unroll
  for (int j = 0; j < 8; ++j) q[j] = (int64_t)((uint64_t)q[j] + (uint64_t)excl);
}

__device__ __forceinline__ void sprintz_tile(const DevChunk &c, uint32_t t, int64_t tile_base, double *out,
                                             double *stage) {
  int64_t q[8];
  sprintz_q(c, t, tile_base, q);
  double v[8];
  const double sc = (double)c.scale;
  if (c.scale) {  // quantized as f64 / scale as f64, correctly rounded via RN(1 / scale)
    const double y = 1.0 / sc;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = div_rn((double)q[j], sc, y);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (double)q[j] / sc;
  }
  tile_store(out + c.rec_off + (uint64_t)t * kTile, v, min(kTile, c.n - t * kTile), stage);
}

__device__ __forceinline__ int64_t sprintz_tile_sum(const DevChunk &c, uint32_t t) {
  const uint32_t i0 = t * kTile + lane_id() * 8;
  int64_t d[8];
  sprintz_deltas(c, i0, d);
  int64_t s = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s = (int64_t)((uint64_t)s + (uint64_t)d[j]);
  return wave_sum_i64(s);
}

// RLE (chunk_reader/run_length.rs:68-91): runs[] u8 and values[] f64 (SoA).
// Tile t covers fields [t*kTile, +kTile); its first output record is the
// tile index entry.  Each lane owns 8 fields; the wave scans run lengths and
// expands step by step (field marks and a max-scan, rle_tile) so that
// output stores are contiguous across lanes.
__device__ __forceinline__ uint32_t rle_runs(const DevChunk &c, uint32_t f0, uint32_t run[8]) {
  const EBI_GLOBAL uint8_t *runs = as_global(reinterpret_cast<const uint8_t *>(c.data));
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    run[j] = (f0 + j < c.fields) ? runs[f0 + j] : 0u;
    s += run[j];
  }
  return s;
}

__device__ __forceinline__ uint64_t rle_tile_sum(const DevChunk &c, uint32_t t) {
  uint32_t run[8];
  const uint32_t s = rle_runs(c, t * kTile + lane_id() * 8, run);
  return wave_sum_u64(s);
}

// starts: this wave's LDS scratch of kTile + 1 entries.

// Expands output steps [s0, s0 + ns) (kTile records each) of field tile t.
__device__ __forceinline__ void rle_tile(const DevChunk &c, uint32_t t, uint32_t s0, uint32_t ns, int64_t tile_start,
                                         double *out, double *stage) {
  const uint32_t lane = lane_id();
  const uint32_t f0 = t * kTile + lane * 8;
  uint32_t run[8], st[8];
  const uint32_t s = rle_runs(c, f0, run);
  const uint32_t incl = wave_incl_scan_u32(s);
  uint32_t pos = incl - s;  // the tile's output position of this lane's first field
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    st[j] = pos;
    pos += run[j];
  }
  const uint32_t total = __shfl(incl, kWave - 1, kWave);
  const EBI_GLOBAL double *vals = as_global(reinterpret_cast<const double *>(c.data + c.values_off)) + (uint64_t)t * kTile;
  double *dst = out + c.rec_off + tile_start;
  uint32_t *mark = reinterpret_cast<uint32_t *>(stage);  // the step's field marks (512 u32), before the stage is used
  // each step expands 512 records: every field marks the first of its
  // records in the step with its index, then record o's field is the
  // largest mark at or before o (fields grow with position): a prefix max
  // over the lane's 8 records and a wave max-scan over the lanes
  const uint32_t end = min(total, (s0 + ns) * kTile);
  for (uint32_t base = s0 * kTile; base < end; base += kTile) {
    uint4 *m4 = reinterpret_cast<uint4 *>(mark) + lane * 2;
    m4[0] = make_uint4(0, 0, 0, 0);
    m4[1] = make_uint4(0, 0, 0, 0);
    wave_sync();
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (run[j] && st[j] < base + kTile && st[j] + run[j] > base) mark[max(st[j], base) - base] = lane * 8 + j;
    wave_sync();
    const uint4 a = m4[0], b = m4[1];
    uint32_t f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 1; j < 8; ++j) f[j] = max(f[j], f[j - 1]);
    uint32_t x = f[7];  // inclusive max-scan of the lanes' last marks
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, kWave);
      if (lane >= (uint32_t)d) x = max(x, y);
    }
    const uint32_t up = __shfl_up(x, 1, kWave);  // every lane takes part in the shuffle
    const uint32_t carry = lane ? up : 0u;
    wave_sync();  // every mark is read before tile_store reuses the stage
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = base + lane * 8 + j < total ? vals[max(f[j], carry)] : 0.0;
    tile_store(dst + base, v, min(kTile, total - base), stage);
  }
}

}  // namespace ebi
