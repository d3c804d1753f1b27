// Cold XOR index: the checkpoints of a Gorilla / Chimp / Chimp128 / Elf /
// ElfOnChimp chunk built in parallel over fixed pieces of its bit stream.
//
// The stream is cut into spans of kSpanBits bits; span i owns the records
// whose first bit lies in [i kSpanBits, (i + 1) kSpanBits) (the chunk's
// last span also the record that would start exactly at the stream's end,
// which can only fail, as it fails in the serial decoder).  Where a span's
// first record starts, and in which parse state, depends on every record
// before it.  The index is found in three steps:
//
//  1. candidates (a lane per kSpansPerLane spans): a guessed walk starts a
//     few spans earlier (xguess_lookback) and runs through the lane's
//     spans; its position and state where it first enters span i is span
//     i's candidate entry.  A misaligned walk of these codes soon falls onto
//     the true one: Chimp, Chimp128, Elf and ElfOnChimp walks adopt the
//     state of every record (a header the codec rejects: skip one bit,
//     restart in the chunk-start state), and the walk itself is then the
//     exact parse of the span from its candidate, so it also gives the
//     span's record count, the XOR of its records' payloads and its last
//     beta*.  Gorilla's state (L, T) changes in rare short bursts and a
//     misaligned 11 header would replace it with noise, so its walk keeps a
//     hinted state (the chunk's, after its first spans).
//  2. walk (a wave per chunk): from the chunk start, span by span: when the
//     true entry equals the span's candidate the candidate's result is
//     taken, else the span is parsed exactly; the running record count,
//     value (prefix XOR: decoder/chunk_reader/gorilla.rs:77-157,
//     chimp.rs:59-125, elf.rs:105-258) and beta* give each span's first
//     record index, the value and beta* before it.  Gorilla's walker first
//     walks kHintSpans spans (the chunk's first burst of state changes),
//     step 1 runs after that in the state it reached, and where the state
//     has moved on the walker guesses the next spans itself.
//  3. emit (a lane per span): each span parsed once more from its true
//     entry, writing the checkpoints of its records and reporting the
//     chunk's first event in record order -- header error, overrun,
//     terminator mismatch, NaN end -- as the serial decoders of codecs.hpp
//     report it.
//
// Result: the same XorCkpt arrays, chunk errors and NaN ends as the serial
// index pass (xor_run / chimp128_index with checkpoints).
#pragma once

#include "codecs.hpp"

namespace ebi {

// A span entry: bit position, parse state (a, b as stored in XorCkpt, 255 =
// unset), whether the record there is the chunk's record 0.  kEntDead: the
// walk that led here ran past the stream, so no record starts in this span.
// kEntGuess: no candidate.
constexpr uint64_t kEntFirst = 1ull << 48;
constexpr uint64_t kEntDead = 1ull << 63;
constexpr uint64_t kEntGuess = ~0ull;

// Parse state at a record boundary (everything but the value and the bit
// position): Gorilla a = L, b = T; Chimp / Chimp128 / ElfOnChimp a = stored
// lead; Elf a = L, b = T (0xffffffff = unset); beta = Elf last beta*
// (INT32_MAX = unset), which only the values depend on.
struct XParse {
  uint32_t a, b;
  int32_t beta;
};

__host__ __device__ __forceinline__ uint32_t ent_u8(uint32_t a) { return a > 64 ? 255u : a; }
__host__ __device__ __forceinline__ uint32_t ent_state(uint64_t e, int sh) {
  const uint32_t v = (uint32_t)(e >> sh) & 0xff;
  return v == 255 ? 0xffffffffu : v;
}
__host__ __device__ __forceinline__ uint64_t ent_pack(uint32_t pos, const XParse &s, bool first) {
  return (uint64_t)pos | ((uint64_t)ent_u8(s.a) << 32) | ((uint64_t)ent_u8(s.b) << 40) | (first ? kEntFirst : 0);
}

__host__ __device__ __forceinline__ XParse xparse_start(uint32_t codec) {
  return codec == kCodecGorilla ? XParse{0u, 0u, INT32_MAX} : XParse{0xffffffffu, 0xffffffffu, INT32_MAX};
}

// Chunk-start entry (record 0 at bit 0).
__host__ __device__ __forceinline__ uint64_t ent_start(uint32_t codec) {
  return ent_pack(0, xparse_start(codec), true);
}

// One record at the window's position: the payload x (value ^= x), the
// bits it uses and the new parse state; kErrCorrupt on a header the
// reference rejects.  The same field logic as the serial decoders
// (gorilla_run, chimp_run, elf_run, elfchimp_run, chimp128_rec).
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ uint32_t xrec(const W &r, bool first, XParse &s, uint64_t &x, uint32_t &used) {
  if (kCodec == kCodecGorilla || kCodec == kCodecChimp || kCodec == kCodecChimp128) {
    if (first) {  // 64 raw bits
      x = r.field(0, 64);
      used = 64;
      return kErrNone;
    }
  }
  if (kCodec == kCodecGorilla) {
    // 0: repeat; 10: reuse (L, T); 11: 5-bit L, 6-bit sig - 1, then sig bits
    const uint32_t c2 = (uint32_t)(r.hi >> 62);
    const bool nz = c2 >= 2, fresh = c2 == 3;
    const uint32_t h = (uint32_t)(r.hi >> 51) & 0x7ff;
    const uint32_t l = fresh ? h >> 6 : s.a;
    const uint32_t sig = fresh ? (h & 63) + 1 : 64 - s.a - s.b;
    if (fresh && l + sig > 64) return kErrCorrupt;
    const uint32_t o = fresh ? 13 : 2;
    x = 0;
    if (nz) {
      const uint32_t t = 64 - l - sig;
      x = r.field(o, sig) << t;
      s.a = l;
      s.b = t;
    }
    used = nz ? o + sig : 1;
    return kErrNone;
  } else if (kCodec == kCodecChimp) {
    const uint32_t f = (uint32_t)(r.hi >> 62);
    const uint32_t h9 = (uint32_t)(r.hi >> 53) & 0x1ff;
    const uint32_t nl = (f == 3 || f == 1) ? lead_repr(h9 >> 6) : s.a;
    if (f == 2 && s.a > 64) return kErrCorrupt;
    uint32_t sig = 64 - nl;
    if (f == 1) {
      sig = (h9 & 63) ? (h9 & 63) : 64;
      if (sig + nl > 64) return kErrCorrupt;
    }
    const uint32_t o = f == 3 ? 5 : f == 1 ? 11 : 2;
    x = 0;
    if (f) {
      x = r.field(o, sig) << (f == 1 ? 64 - sig - nl : 0);
      s.a = nl;
    }
    used = f ? o + sig : 2;
    return kErrNone;
  } else if (kCodec == kCodecChimp128) {
    uint32_t f, idx;
    if (!chimp128_rec(r, s.a, f, idx, x, used)) return kErrCorrupt;
    return kErrNone;
  } else if (kCodec == kCodecElf) {
    const uint32_t p2 = (uint32_t)(r.hi >> 62);
    const uint32_t pre = (p2 & 2) ? (p2 == 2 ? 2u : 6u) : 1u;
    if (p2 == 3) s.beta = (int32_t)((r.hi >> 58) & 15);
    if (first) {  // ElfXor first value: 7-bit trailing zeros, then 63 - tz bits
      const uint32_t tz = (uint32_t)r.field(pre, 7);
      if (tz < 64) {
        const uint32_t len = 63 - tz;
        const uint64_t m = len ? r.field(pre + 7, len) : 0;
        x = ((m << 1) | 1) << tz;
        used = pre + 7 + len;
      } else {
        x = 0;
        used = pre + 7;
      }
      return kErrNone;
    }
    const uint32_t f = (uint32_t)r.field(pre, 2);
    const uint32_t h9 = (uint32_t)r.field(pre + 2, 9), h7 = h9 >> 2;
    uint32_t nL = s.a, nT = s.b, len = 0, off = pre + 2;
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
      if ((uint64_t)s.a + s.b > 64) return kErrCorrupt;
      len = 64 - s.a - s.b;
    }
    const uint64_t m = len ? r.field(off, len) : 0;
    used = f == 1 ? pre + 2 : off + len;
    x = f == 1 ? 0 : f == 0 ? m << s.b : ((m << 1) | 1) << nT;
    s.a = nL;
    s.b = nT;
    return kErrNone;
  } else {  // kCodecElfOnChimp
    const uint32_t p2 = (uint32_t)(r.hi >> 62);
    const uint32_t pre = (p2 & 2) ? (p2 == 2 ? 2u : 6u) : 1u;
    if (p2 == 3) s.beta = (int32_t)((r.hi >> 58) & 15);
    if (first) {
      x = r.field(pre, 64);
      used = pre + 64;
      return kErrNone;
    }
    const uint32_t f = (uint32_t)r.field(pre, 2);
    const uint32_t h9 = (uint32_t)r.field(pre + 2, 9);
    const uint32_t nl = (f == 3 || f == 1) ? lead_repr(h9 >> 6) : s.a;
    if (f == 2 && s.a > 64) return kErrCorrupt;
    uint32_t sig = 64 - nl;
    if (f == 1) {
      sig = (h9 & 63) ? (h9 & 63) : 64;
      if (sig + nl > 64) return kErrCorrupt;
    }
    const uint32_t o = pre + (f == 3 ? 5u : f == 1 ? 11u : 2u);
    x = 0;
    if (f) {
      x = r.field(o, sig) << (f == 1 ? 64 - sig - nl : 0);
      s.a = nl;
    }
    used = f ? o + sig : pre + 2;
    return kErrNone;
  }
}

// Span i of a chunk of `bits` bits ends at span_end (exclusive; the last
// span, i + 1 == span_count, also owns a record starting at `bits`).
__host__ __device__ __forceinline__ uint32_t span_end(uint32_t i, uint32_t bits) {
  const uint64_t e = (uint64_t)(i + 1) * kSpanBits;
  return e < bits ? (uint32_t)e : bits;
}

// Guessed walk from the window's position in state s to `end` (on a header
// the codec rejects: drop one bit, restart in the chunk-start state): the
// entry it reaches.  Continues the exact walk of a span after an error.
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ uint64_t guess_walk(W &r, XParse s, uint32_t end, bool incl, uint32_t bits) {
  while (r.pos < end || (incl && r.pos == end)) {
    uint64_t x;
    uint32_t used;
    if (xrec<kCodec>(r, false, s, x, used)) {
      used = 1;
      s = xparse_start(kCodec);
    }
    r.consume(used);
    r.refill();
    if (r.pos > bits) return kEntDead;
  }
  return ent_pack(r.pos, s, false);
}

// Exact walk of a span's records from the window's position and state s:
// records starting before `end` (and at `end` when incl).  Vis::before(pos,
// s) is called ahead of each record and may stop the walk (false);
// Vis::after(x, s) after it.  Returns the error that ended the walk
// (kErrCorrupt / kErrOverrun; the failing record is not counted) or
// kErrNone.
template <uint32_t kCodec, class W, class Vis>
__host__ __device__ __forceinline__ uint32_t span_walk(W &r, XParse &s, bool &first, uint32_t end, bool incl,
                                                   uint32_t bits, Vis &vis) {
  while (r.pos < end || (incl && r.pos == end)) {
    if (!vis.before(r.pos, s)) return kErrNone;
    uint64_t x;
    uint32_t used;
    const uint32_t e = xrec<kCodec>(r, first, s, x, used);
    if (e) return e;
    r.consume(used);
    r.refill();
    if (r.pos > bits) return kErrOverrun;
    first = false;
    vis.after(x, s);
  }
  return kErrNone;
}

struct CountVis {
  uint32_t cnt = 0;
  uint64_t xr = 0;
  int32_t beta0;
  __host__ __device__ __forceinline__ bool before(uint32_t, const XParse &) const { return true; }
  __host__ __device__ __forceinline__ void after(uint64_t x, const XParse &) {
    xr ^= x;
    ++cnt;
  }
};

// Step 2 for one span from entry `ent` (not kEntGuess).  A walk that fails
// (a header error, or a record running past the stream) records the error;
// after a header error the exit comes from a guessed walk on from the next
// bit, so that a wrong entry does not leave every later span without one.
// (When the entry is the true one the chunk has failed there, and the scan
// drops every later span.)
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ SpanSum span_settle(const uint8_t *data, uint64_t ent, uint32_t end, bool incl,
                                                   uint32_t bits) {
  if (ent & kEntDead) return SpanSum{kEntDead, 0, 0, -1, kErrOverrun};
  W r;
  const uint32_t pos = (uint32_t)ent;
  r.init(data, pos);
  XParse s{ent_state(ent, 32), ent_state(ent, 40), INT32_MAX};
  bool first = (ent & kEntFirst) != 0;
  CountVis v;
  const uint32_t e = span_walk<kCodec>(r, s, first, end, incl, bits, v);
  const int16_t lb = (int16_t)(s.beta == INT32_MAX ? -1 : s.beta);
  if (!e) return SpanSum{ent_pack(r.pos, s, first), v.xr, v.cnt, lb, 0};
  if (e == kErrOverrun) return SpanSum{kEntDead, v.xr, v.cnt, lb, (uint16_t)e};
  r.consume(1);
  r.refill();
  const uint64_t ex = r.pos > bits ? kEntDead : guess_walk<kCodec>(r, xparse_start(kCodec), end, incl, bits);
  return SpanSum{ex, v.xr, v.cnt, lb, (uint16_t)e};
}

// Step 4's visitor: checkpoints every `every` records (kSeg; Chimp128
// kSeg128, parse state only, through its terminator record n), and the
// span's first event in record order.
struct EmitVis {
  uint32_t g;     // index of the next record
  uint32_t n;     // footer record count (the terminator's index)
  uint64_t v;     // value before the next record
  int32_t beta;   // beta* before the next record (INT32_MAX unset)
  XorCkpt *ck;
  uint32_t codec;
  uint32_t ev_rec, ev_code;  // first event (ev_code 0: none yet)
  __host__ __device__ __forceinline__ bool before(uint32_t pos, const XParse &s) {
    if (g > n || ev_code) return false;
    if (codec == kCodecChimp128) {
      if (g % kSeg128 == 0) ck[g / kSeg128] = XorCkpt{0, pos, (uint8_t)ent_u8(s.a), 0, -1, 0};
    } else if (g < n && g % kSeg == 0) {
      const bool lt = codec == kCodecGorilla;
      ck[g / kSeg] = XorCkpt{v, pos, (uint8_t)(lt ? s.a : ent_u8(s.a)),
                             (uint8_t)(codec == kCodecChimp || codec == kCodecElfOnChimp ? 0 : lt ? s.b : ent_u8(s.b)),
                             (int8_t)(beta == INT32_MAX ? -1 : beta), 0};
    }
    return true;
  }
  __host__ __device__ __forceinline__ void after(uint64_t x, const XParse &s) {
    v ^= x;
    beta = s.beta;
    if (codec != kCodecChimp128) {
      // Elf keeps its stored value when the new one would be NaN (that
      // record ends the chunk), so the XOR form holds up to the first NaN
      if (g == n) {
        if (v != kNan) {
          ev_rec = g;
          ev_code = kErrCount;
        }
      } else if (v == kNan) {
        ev_rec = g;
        ev_code = kNanEnd;
      }
    }
    ++g;
  }
};

// Step 4 for one span: entry `ent`, its first record index g0, the value
// and beta* before it.  Returns the span's first event as (record << 8 |
// code), or ~0ull.
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ uint64_t span_emit(const uint8_t *data, uint64_t ent, uint32_t g0, uint64_t v0,
                                                  int32_t beta0, uint32_t end, bool incl, uint32_t bits, uint32_t n,
                                                  XorCkpt *ck) {
  if ((ent & kEntDead) || g0 > n) return ~0ull;
  W r;
  r.init(data, (uint32_t)ent);
  XParse s{ent_state(ent, 32), ent_state(ent, 40), beta0};
  bool first = (ent & kEntFirst) != 0;
  EmitVis v{g0, n, v0, beta0, ck, kCodec, 0, 0};
  const uint32_t e = span_walk<kCodec>(r, s, first, end, incl, bits, v);
  if (v.ev_code) return ((uint64_t)v.ev_rec << 8) | v.ev_code;
  if (e && v.g <= n) return ((uint64_t)v.g << 8) | e;
  return ~0ull;
}

// ------------------------------------------------------------ step 1

// Spans a codec's guessed walk runs before the first span it gives a
// candidate for (the distances over which misaligned walks were measured
// to fall onto the true one on the four column generators: Elf within one
// span, Chimp128 two, Chimp and a Gorilla walk in the true state three).
__host__ __device__ __forceinline__ uint32_t xguess_lookback(uint32_t codec) {
  return codec == kCodecChimp || codec == kCodecGorilla ? 3u : (codec == kCodecElf || codec == kCodecElfOnChimp) ? 1u : 2u;
}

// Header errors inside the spans it gives candidates for after which a
// hinted walk gives up: a walk in the true state meets its errors before it
// falls onto the record boundaries, mostly in the spans it runs through
// first; one in a state the stream has left meets one every few records.
constexpr uint32_t kHintedErrs = 4;

// Candidates of spans [j, j1) of a chunk of ns spans from one walk started
// at span a: cand[i] (the walk's position and state where it enters span i)
// and res[i] = span_settle(cand[i]).  An adopting walk through span i is
// that exact parse (same records, same error, and after an error the same
// guessed continuation).  A hinted walk (fixed: Gorilla from a span after
// the first; hint = the state, a packed entry whose position is ignored)
// keeps the hinted state, so it is the exact parse of a span as long as no
// record in it changes the state (an 11 header with a new L, T) or fails;
// a span where one does, and every span after more than max_errs errors
// in the spans it gives candidates for, gets no candidate (kEntGuess).  A walk from bit 0 is the exact walk from the chunk
// start; so is one from a true entry `from` (with fixed false).
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ void cand_walk(const uint8_t *data, uint32_t bits, uint32_t ns, uint32_t j,
                                               uint32_t j1, uint32_t a, bool fixed, uint64_t hint, uint64_t *cand,
                                               SpanSum *res, uint64_t from = kEntGuess,
                                               uint32_t max_errs = kHintedErrs) {
  const XParse s0 = fixed ? XParse{ent_state(hint, 32), ent_state(hint, 40), INT32_MAX} : xparse_start(kCodec);
  W r;
  XParse s = s0;
  bool first = a == 0;
  if (from != kEntGuess) {  // the exact walk from a true entry
    r.init(data, (uint32_t)from);
    s.a = ent_state(from, 32);
    s.b = ent_state(from, 40);
    first = (from & kEntFirst) != 0;
  } else {
    r.init(data, a * kSpanBits);
  }
  uint32_t nxt = j;  // the span whose entry comes next
  uint32_t errs = 0;
  bool open = false, div = false;
  SpanSum acc{0, 0, 0, -1, 0};
  for (;;) {
    const uint32_t p = r.pos;
    if (nxt < ns && p >= (uint64_t)nxt * kSpanBits) {  // enters span nxt (a record is shorter than a span)
      const uint64_t e = ent_pack(p, s, first);
      if (open) {
        acc.exit = e;
        if (div) cand[nxt - 1] = kEntGuess; else res[nxt - 1] = acc;
      }
      if (nxt == j1) return;
      cand[nxt] = e;
      acc = SpanSum{0, 0, 0, -1, 0};
      open = true;
      div = false;
      ++nxt;
    }
    uint64_t x;
    uint32_t used;
    XParse t = s;
    t.beta = INT32_MAX;  // set by the record when it carries a beta*
    const uint32_t err = xrec<kCodec>(r, first, t, x, used);
    if (err) {
      if (open && !acc.err) acc.err = (uint16_t)err;
      if (fixed && open && ++errs > max_errs) {  // a state the stream has left: no candidates from here
        for (uint32_t i = open ? nxt - 1 : nxt; i < j1; ++i) cand[i] = kEntGuess;
        return;
      }
      div |= fixed;  // the exact parse goes on in the chunk-start state
      used = 1;
      s = fixed ? s0 : xparse_start(kCodec);
    } else {
      if (open && !acc.err) {
        acc.xr ^= x;
        ++acc.cnt;
        if (t.beta != INT32_MAX) acc.lbeta = (int16_t)t.beta;
      }
      if (!fixed) {
        s.a = t.a;
        s.b = t.b;
      }
      div |= s.a != t.a || s.b != t.b;
    }
    first = false;
    r.consume(used);
    r.refill();
    if (r.pos > bits) {  // a record runs past the stream: no later span has an entry
      if (open) {
        if (!acc.err) acc.err = kErrOverrun;
        acc.exit = kEntDead;
        if (div) cand[nxt - 1] = kEntGuess; else res[nxt - 1] = acc;
      }
      for (; nxt < j1; ++nxt) {
        cand[nxt] = kEntDead;
        res[nxt] = SpanSum{kEntDead, 0, 0, -1, kErrOverrun};
      }
      return;
    }
  }
}

template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ void span_cands(const uint8_t *data, uint32_t bits, uint32_t ns, uint32_t j,
                                                uint32_t j1, uint64_t hint, uint64_t *cand, SpanSum *res) {
  const uint32_t lb = xguess_lookback(kCodec);
  const uint32_t a = j > lb ? j - lb : 0;
  cand_walk<kCodec, W>(data, bits, ns, j, j1, a, kCodec == kCodecGorilla && a > 0, hint, cand, res);
}

// ------------------------------------------------------------ step 2
// Candidates of spans [j, j1) of a Gorilla walker's block that starts at
// span b0 (true entry `from`): one walk in `from`'s state from
// kBlockLookback spans before j, or the exact walk from `from` when that
// would start before b0.  A block covers kWaveLanes * kBlockSpansPerLane
// spans, kBlockSpansPerLane per walk: after a state change the stream
// usually keeps the new state to the chunk's end, and every block boundary
// is a miss (its span was guessed in the old state), so longer blocks save
// boundaries at the price of longer walks.
constexpr uint32_t kBlockLookback = 2;
constexpr uint32_t kBlockSpansPerLane = 4;
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ void block_cands(const uint8_t *data, uint32_t bits, uint32_t ns, uint32_t j,
                                                 uint32_t j1, uint32_t b0, uint64_t from, uint64_t *cand,
                                                 SpanSum *res) {
  if (j < b0 + kBlockLookback)
    cand_walk<kCodec, W>(data, bits, ns, j, j1, b0, false, 0, cand, res, from);
  else
    cand_walk<kCodec, W>(data, bits, ns, j, j1, j - kBlockLookback, true, from, cand, res);
}

// Step 2 for one chunk from w.frontier up to span `stop` (Gorilla first
// walks to span kHintSpans, whose state becomes the hint of its candidate
// walks).  Gorilla (blocks): after a miss whose exact parse ends in a state
// other than the one the next span's candidate was guessed in, the walker
// guesses the next kWaveLanes * kBlockSpansPerLane spans itself in the new
// state (block_cands).
// After a span whose true walk fails, the later spans get kEntDead entries.
template <uint32_t kCodec, class W>
__host__ __device__ __forceinline__ void chunk_walk(const uint8_t *data, uint32_t bits, uint32_t ns, ChunkWalk &w,
                                                uint64_t *cand, SpanSum *res, SpanPick *pick, uint32_t stop,
                                                bool blocks) {
  const uint32_t end = min(ns, stop), chint = (uint32_t)(w.cur >> 32) & 0xffff;
  uint32_t bend = 0, bhint = 0;  // spans before bend were guessed in state bhint, the others in chint
  for (uint32_t i = w.frontier; i < end; ++i) {
    const bool hit = cand[i] == w.cur;
    const SpanSum s = hit ? res[i] : span_settle<kCodec, W>(data, w.cur, span_end(i, bits), i + 1 == ns, bits);
    pick[i] = SpanPick{w.cur, w.v, w.g, w.beta};
    if (s.err) {
      for (uint32_t k = i + 1; k < ns; ++k) pick[k] = SpanPick{kEntDead, 0, 0, INT32_MAX};
      w.frontier = ns;
      return;
    }
    w.g += s.cnt;
    w.v ^= s.xr;
    if (s.lbeta >= 0) w.beta = s.lbeta;
    w.cur = s.exit;
    const uint32_t st = (uint32_t)(w.cur >> 32) & 0xffff, nx = i + 1;
    if (blocks && !hit && nx < end && st != (nx < bend ? bhint : chint)) {
      bend = min(end, nx + kWaveLanes * kBlockSpansPerLane);
      bhint = st;
      for (uint32_t k = nx; k < bend; k += kBlockSpansPerLane)
        block_cands<kCodec, W>(data, bits, ns, k, min(k + kBlockSpansPerLane, bend), nx, w.cur, cand, res);
    }
  }
  w.frontier = end;
}

// Dispatch helpers on the codec.
template <class W>
__host__ __device__ __forceinline__ SpanSum span_settle_any(uint32_t codec, const uint8_t *data, uint64_t ent,
                                                       uint32_t end, bool incl, uint32_t bits) {
  switch (codec) {
    case kCodecGorilla: return span_settle<kCodecGorilla, W>(data, ent, end, incl, bits);
    case kCodecChimp: return span_settle<kCodecChimp, W>(data, ent, end, incl, bits);
    case kCodecChimp128: return span_settle<kCodecChimp128, W>(data, ent, end, incl, bits);
    case kCodecElf: return span_settle<kCodecElf, W>(data, ent, end, incl, bits);
    default: return span_settle<kCodecElfOnChimp, W>(data, ent, end, incl, bits);
  }
}
template <class W>
__host__ __device__ __forceinline__ void span_cands_any(uint32_t codec, const uint8_t *data, uint32_t bits, uint32_t ns,
                                                    uint32_t j, uint32_t j1, uint64_t hint, uint64_t *cand,
                                                    SpanSum *res) {
  switch (codec) {
    case kCodecGorilla: return span_cands<kCodecGorilla, W>(data, bits, ns, j, j1, hint, cand, res);
    case kCodecChimp: return span_cands<kCodecChimp, W>(data, bits, ns, j, j1, hint, cand, res);
    case kCodecChimp128: return span_cands<kCodecChimp128, W>(data, bits, ns, j, j1, hint, cand, res);
    case kCodecElf: return span_cands<kCodecElf, W>(data, bits, ns, j, j1, hint, cand, res);
    default: return span_cands<kCodecElfOnChimp, W>(data, bits, ns, j, j1, hint, cand, res);
  }
}
template <class W>
__host__ __device__ __forceinline__ void chunk_walk_any(uint32_t codec, const uint8_t *data, uint32_t bits,
                                                    uint32_t ns, ChunkWalk &w, uint64_t *cand, SpanSum *res,
                                                    SpanPick *pick, uint32_t stop, bool blocks) {
  switch (codec) {
    case kCodecGorilla: return chunk_walk<kCodecGorilla, W>(data, bits, ns, w, cand, res, pick, stop, blocks);
    case kCodecChimp: return chunk_walk<kCodecChimp, W>(data, bits, ns, w, cand, res, pick, stop, false);
    case kCodecChimp128: return chunk_walk<kCodecChimp128, W>(data, bits, ns, w, cand, res, pick, stop, false);
    case kCodecElf: return chunk_walk<kCodecElf, W>(data, bits, ns, w, cand, res, pick, stop, false);
    default: return chunk_walk<kCodecElfOnChimp, W>(data, bits, ns, w, cand, res, pick, stop, false);
  }
}
template <class W>
__host__ __device__ __forceinline__ uint64_t span_emit_any(uint32_t codec, const uint8_t *data, uint64_t ent,
                                                      uint32_t g0, uint64_t v0, int32_t beta0, uint32_t end,
                                                      bool incl, uint32_t bits, uint32_t n, XorCkpt *ck) {
  switch (codec) {
    case kCodecGorilla: return span_emit<kCodecGorilla, W>(data, ent, g0, v0, beta0, end, incl, bits, n, ck);
    case kCodecChimp: return span_emit<kCodecChimp, W>(data, ent, g0, v0, beta0, end, incl, bits, n, ck);
    case kCodecChimp128: return span_emit<kCodecChimp128, W>(data, ent, g0, v0, beta0, end, incl, bits, n, ck);
    case kCodecElf: return span_emit<kCodecElf, W>(data, ent, g0, v0, beta0, end, incl, bits, n, ck);
    default: return span_emit<kCodecElfOnChimp, W>(data, ent, g0, v0, beta0, end, incl, bits, n, ck);
  }
}

}  // namespace ebi
