// Decode kernels of libebi_hip (gfx950) and their host launchers.
//
// Launch shapes:
//  * k_decode   -- persistent, 256 threads = 4 independent waves; each wave
//                  takes tile jobs (up to four 512-record tiles of a tile
//                  codec) from the job space of every column in the launch;
//  * k_decode_seg -- the same over checkpoint-segment jobs (64 segments of
//                  an XOR chunk, one per lane);
//  * k_serial   -- 64 threads; lane l walks chunk (job.chunk + l) serially
//                  (Gorilla, Chimp, Elf); also the load-time index pass that
//                  writes XOR checkpoints;
//  * k_chimp128 -- 64 threads with a 64 KiB LDS ring (128 slots per lane);
//  * k_tile_sums / k_tile_scan -- the load-time tile index of DeltaSprintz
//                  (prefix q per tile) and RLE (first output record per tile).
#include "codecs.hpp"
#include "launch.hpp"
#include "query.hpp"

namespace ebi {

// Column c with prefix[c] <= j < prefix[c + 1] (wave-uniform binary search).
__device__ __forceinline__ uint32_t find_col(const uint32_t *prefix, uint32_t n_cols, uint32_t j) {
  uint32_t lo = 0, hi = n_cols;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= j) lo = mid; else hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// The column of launch-wide job j, remembered across a persistent wave's
// jobs: the wave's successive jobs (a grid stride apart) mostly fall in the
// same column, so the binary search over the prefix -- a chain of dependent
// loads at every job start -- runs only when j leaves that column's range.
// All fields wave-uniform.  Also used over the plans of a query group.
struct ColCursor {
  uint32_t ci = 0, lo = 1, hi = 0;  // column ci holds the jobs [lo, hi); empty at first
  __device__ __forceinline__ uint32_t at(const uint32_t *prefix, uint32_t n_cols, uint32_t j) {
    if (j < lo || j >= hi) {
      ci = find_col(prefix, n_cols, j);
      lo = __builtin_amdgcn_readfirstlane(prefix[ci]);
      hi = __builtin_amdgcn_readfirstlane(prefix[ci + 1]);
    }
    return ci;
  }
};

// State of a lane with no segment in a lockstep wave: every field defined
// (zeros), position 0, so refill() and field() compute on known values.
__device__ __forceinline__ void dead_window(BitWindow &r) {
  r.p = nullptr;
  r.c0 = r.c1 = r.n0 = r.n1 = 0;
  r.q[0] = make_uint4(0, 0, 0, 0);
  r.b = 0;
  r.stride = 0;
  r.hi = r.lo = 0;
  r.pos = 0;
}
__device__ __forceinline__ void dead_state(XorState &s) {
  dead_window(s.r);
  s.v = 0;
  s.a = s.b = 0;
  s.beta = 0;
  s.end = 0;
}

// Checkpoint segments first .. first + 63 of chunk c (Gorilla, Chimp or
// Elf), segment first + l on lane l, decoded in lockstep runs of kRun
// records from the checkpoints into `out`: out.row(i, k, v, act) takes each
// record, out.run_end(k0) follows each run.  The loop over runs is
// wave-uniform: lanes without a segment (past the chunk's end) stay in it.
// The chunks reaching here were parsed whole and accepted by the load-time
// index (a rejected chunk fails the call before any launch), so the records
// go through the codec loops without their per-record checks; record 0's
// check remains.
template <uint32_t kCodec, class Out>
__device__ __forceinline__ uint32_t segment_wave(const ColView &cv, const DevChunk &c, uint32_t first, uint32_t row,
                                                const double *elf_tab, Out &out) {
  const uint32_t i0 = (first + lane_id()) * kSeg;
  const bool live = i0 < c.n;
  const uint32_t i1 = live ? min(i0 + kSeg, c.n) : 0;
  XorState s;
  if (live) {
    const XorCkpt k = reinterpret_cast<const XorCkpt *>(c.xidx)[first + lane_id()];
    s.r.init_at(cv.ilv + (size_t)row * kIlvRowBlocks + lane_id(), k.bitpos, kIlvRowBlocks);
    xor_state_set(s, kCodec, k);
  }
  const uint32_t bits = c.size * 8u;
  // a lane without a segment has an empty, fully defined window that it
  // never advances (it takes part in no record, but the lockstep loop still
  // runs refill() and the field extraction on it)
  if (!live) dead_state(s);
  const uint32_t last = live && i1 == c.n ? c.n + 1 : i1;  // the chunk's last segment parses the terminator
  uint32_t from = i0, e = 0;
  if (live && i0 == 0) {
    e = kCodec == kCodecElf          ? elf_seg_first(s, bits, out, elf_tab)
        : kCodec == kCodecElfOnChimp ? elfchimp_seg_first(s, bits, out, elf_tab)
                                     : xor_seg_first(s, bits, out);
    from = 1;
  }
#pragma unroll 1  // one copy of the codec loop per codec (instruction cache)
  for (uint32_t k0 = 0; k0 < kSeg; k0 += kRun) {
    const uint32_t a = i0 + k0;
    // kRun records, and one more when a lane's terminator follows them
    const bool term_next = last == a + kRun + 1;
    const uint32_t cnt = kRun + (__ballot(term_next) != 0);
    const uint32_t lim = term_next ? last : min(last, a + kRun);
    if (kCodec == kCodecGorilla)
      e = gorilla_seg<false>(s, bits, c.n, a, cnt, from, lim, out, e);
    else if (kCodec == kCodecChimp)
      e = chimp_seg<false>(s, bits, c.n, a, cnt, from, lim, out, e);
    else if (kCodec == kCodecElfOnChimp)
      e = elfchimp_seg<false>(s, bits, c.n, a, cnt, from, lim, out, e, elf_tab);
    else
      e = elf_seg<false>(s, bits, c.n, a, cnt, from, lim, out, e, elf_tab);
    from = max(from, lim);  // a terminator taken as record a + kRun is not parsed again
    out.run_end(k0);
  }
  return e;
}

// Decode output of segment_wave: the wave's LDS stage, written out after
// each run (stage_flush) to the job's records in cv.out.
struct StageOut {
  StageSink sink;
  double *dst;      // record 0 of segment `first`
  uint32_t n_job;   // records of the chunk from segment `first` on
  __device__ __forceinline__ void row(uint32_t i, uint32_t k, double v, bool act) { sink.row(i, k, v, act); }
  __device__ __forceinline__ void run_end(uint32_t k0) { stage_flush(sink.stage, dst, k0, n_job); }
};

// Chimp128 segment job: segments first .. first + 63 of chunk c (kSeg128
// records each; records 0..n, n the terminator), segment first + l on lane
// l, one job = kJob128 records.  The wave parses every segment in one
// lockstep run (payloads into the stage, row k = record i0 + k; source
// distances into `dists`, job record t at byte t), then resolves the job's records in
// record order in steps of 64 (lane = record): sources inside the step by
// pointer jumping over lane shuffles (x ^= x[src], src = src[src] while the
// source lies in the step; sources only point back, so this ends after at
// most six rounds), sources in earlier steps from the stage (each step's
// values replace its payloads), sources before the job from its ring
// checkpoint (cv.c128_ring + 128 ring_idx).  out.put(i, v, base) takes
// record i's value (base: the step's first record; every lane calls it).
// With kChecked, record n must be the NaN end marker and an earlier NaN is
// a count error (a NaN can only come from flags 01 / 10 / 11; the ring never
// holds one, as the reference stops at the first), and the parse checks
// each record: the load-time ring build, which is where a Chimp128 chunk is
// validated (with the index's parse).  The decode and the fused queries run
// on accepted chunks only and pass kChecked false.
template <bool kChecked, class Out>
__device__ __forceinline__ uint32_t c128_job(const ColView &cv, const DevChunk &c, uint32_t first, uint32_t ring_idx,
                                            double *stage, uint8_t *dists, Out &out) {
  const uint32_t lane = lane_id();
  const uint32_t total = c.n + 1;
  const uint32_t i0 = (first + lane) * kSeg128;
  const bool live = i0 < total;
  BitWindow r;
  uint32_t lead = 0xffffffffu;
  if (live) {
    const XorCkpt k = reinterpret_cast<const XorCkpt *>(c.xidx)[first + lane];
    r.init(reinterpret_cast<const uint8_t *>(c.data), k.bitpos);
    lead = k.a == 255 ? 0xffffffffu : k.a;
  } else {  // an empty, fully defined window that is never advanced
    dead_window(r);
  }
  const StageSink sink{stage, dists};
  const uint32_t bits = c.size * 8u;
  uint32_t from = i0, e = 0;
  uint64_t px = 0;
  uint32_t pd = 0;
  if (live && i0 == 0) {
    e = chimp128_seg_first(r, bits, sink, px, pd);
    from = 1;
  }
  e = chimp128_parse_seg<kChecked>(r, lead, bits, i0, kSeg128, from, live ? min(i0 + kSeg128, total) : 0u, sink, e,
                                   px, pd, live && i0 == 0);
  wave_sync();  // the stage holds every payload of the job
  const uint32_t j0 = first * kSeg128;  // the job's first record
  const EBI_GLOBAL uint64_t *ring = as_global(cv.c128_ring + 128ull * ring_idx);
#pragma unroll 1
  for (uint32_t s = 0; s < kJob128 / kWave; ++s) {
    const uint32_t base = j0 + s * kWave;
    if (base >= total) break;
    const uint32_t t = s * kWave + lane;  // record j0 + t: row t % 16 of lane t / 16's segment
    const uint32_t i = j0 + t;
    const uint32_t sg = t / kSeg128, kk = t % kSeg128;
    double *slot = stage + kk * kWave + (sg ^ stage_swz(kk));
    uint64_t v = as_u64(*slot);
    const uint32_t d = i < total ? dists[t] : 0u;
    int32_t p = d ? (int32_t)(i - d) : -1;  // source record, -1: none (value 0)
    while (__any(p >= (int32_t)base)) {
      const int32_t q = (p - (int32_t)base) & (kWave - 1);  // source lane (used only when in the step)
      const uint64_t vq = __shfl(v, q);
      const int32_t pq = __shfl(p, q);
      const bool in = p >= (int32_t)base;
      v = in ? v ^ vq : v;
      p = in ? pq : p;
    }
    if (p >= (int32_t)j0) {  // an earlier step of this job: its value replaced the payload
      const uint32_t tp = (uint32_t)p - j0;
      v ^= as_u64(stage[(tp % kSeg128) * kWave + ((tp / kSeg128) ^ stage_swz(tp % kSeg128))]);
    } else if (p >= 0) {  // before the job: its ring checkpoint
      v ^= ring[(uint32_t)p + 128 - j0];
    }
    if (i < total) *slot = as_f64(v);
    wave_sync();  // this step's values are in the stage before the next step reads them
    if (kChecked) {
      if (i < c.n) {
        if (v == kNan && !Out::kNanEnds) e |= kErrCount;
      } else if (i == c.n && v != kNan) {
        e |= kErrCount;
      }
    }
    out.put(i, v, base);
  }
  return e;
}

// Decode output of c128_job: record i < n to cv.out (a step's 64 records are
// one 512-byte store).
struct C128Out {
  static constexpr bool kNanEnds = false;  // chunks end at their first NaN since load
  double *dst;  // record 0 of the chunk
  uint32_t n;
  __device__ __forceinline__ void put(uint32_t i, uint64_t v, uint32_t) {
    if (i < n) st_one(dst + i, as_f64(v));
  }
};

__device__ __forceinline__ uint32_t segment_job(const ColView &cv, uint32_t ci, uint32_t first, uint32_t row,
                                               double *stage, uint8_t *dists, const double *elf_tab) {
  const DevChunk c = cv.chunks[ci];
  StageOut out{StageSink{stage, nullptr}, cv.out + c.rec_off + (size_t)first * kSeg,
               first * kSeg < c.n ? c.n - first * kSeg : 0u};
  switch (c.codec) {
    case kCodecGorilla: return segment_wave<kCodecGorilla>(cv, c, first, row, elf_tab, out);
    case kCodecChimp: return segment_wave<kCodecChimp>(cv, c, first, row, elf_tab, out);
    case kCodecElf: return segment_wave<kCodecElf>(cv, c, first, row, elf_tab, out);
    case kCodecElfOnChimp: return segment_wave<kCodecElfOnChimp>(cv, c, first, row, elf_tab, out);
    case kCodecChimp128: {
      C128Out o{cv.out + c.rec_off, c.n};
      return c128_job<false>(cv, c, first, row, stage, dists, o);
    }
    default: return kErrCorrupt;
  }
}

// Tile job jl of column cv (Uncompressed, BUFF, DeltaSprintz, RLE): kCodec
// is the codec of every chunk of the launch (its tile loop alone is compiled
// in), or kAnyCodec (each chunk's own).  stage: the wave's store stage
// (kStageStride * kWave doubles; an RLE step's field marks before that).
constexpr uint32_t kAnyCodec = 0xffu;
template <uint32_t kCodec>
__device__ __forceinline__ void tile_job_of(const ColView &cv, uint32_t jl, double *stage, uint32_t *err) {
  // jb.kind consecutive tiles from jb.tile; RLE: jb.kind output steps
  // from step jb.tile >> 24 of field tile jb.tile & 0xffffff
  const Job jb = cv.tile_jobs[jl];
  const DevChunk c = cv.chunks[jb.chunk];
  const uint32_t codec = kCodec == kAnyCodec ? c.codec : kCodec;
  if (cv.chunk_err && cv.chunk_err[jb.chunk]) {
    // chunk rejected at load: nothing to decode
  } else if (codec != c.codec) {
    if (lane_id() == 0) atomicOr(err, (uint32_t)kErrCorrupt);
  } else if (codec == kCodecRLE) {
    const uint32_t t = jb.tile & 0xffffffu;
    rle_tile(c, t, jb.tile >> 24, jb.kind, cv.tile_index[c.tile0 + t], cv.out, stage);
  } else {
    for (uint32_t t = jb.tile; t < jb.tile + jb.kind; ++t) {
      switch (codec) {
        case kCodecPlain: plain_tile(c, t, cv.out); break;
        case kCodecBUFF: buff_tile(c, t, cv.out, stage, cv.precision); break;
        case kCodecSprintz: sprintz_tile(c, t, cv.tile_index[c.tile0 + t], cv.out, stage); break;
        default:
          if (lane_id() == 0) atomicOr(err, (uint32_t)kErrCorrupt);
          break;
      }
    }
  }
}

// Tile job j of the launch (the columns' tile_jobs lists concatenated).
__device__ __forceinline__ void tile_job(const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t j,
                                         double *stage, uint32_t *err) {
  const uint32_t ci = find_col(prefix, n_cols, j);
  tile_job_of<kAnyCodec>(cols[ci], j - prefix[ci], stage, err);
}

// Segment job jl of column cv (64 checkpoint segments of a Gorilla / Chimp /
// Elf chunk, or a Chimp128 job).
__device__ __forceinline__ void seg_job_of(const ColView &cv, uint32_t jl, double *stage, uint8_t *dists,
                                           const double *elf_tab, uint32_t *err) {
  const Job jb = cv.seg_jobs[jl];
  if (!cv.chunk_err || cv.chunk_err[jb.chunk] == 0) {
    const uint32_t e = segment_job(cv, jb.chunk, jb.tile, jb.col, stage, dists, elf_tab);
    if (e) atomicOr(err, e);
  }
}

// Persistent decode over the tile jobs of the columns of one tile codec
// among n_cols columns (uneven chunks in one launch; one launch per codec
// present).  Wave w takes jobs w, w + W, w + 2W, ... of the launch's whole
// job list (W = waves in the grid): a static interleave, so neighbouring
// jobs (similar cost) spread over all waves and no job counter is
// contended; a wave that reaches a column of another codec steps past its
// jobs at once.  One kernel per codec so that each gets the waves per SIMD
// its own registers allow (TileWaves: the memory-bound copy at 8, BUFF at 5,
// RLE at 6, DeltaSprintz's wave scans at 5), to keep more loads in flight.
template <uint32_t kCodec>
struct TileWaves {
  static constexpr int value = kCodec == kCodecSprintz ? 5 : kCodec == kCodecRLE ? 6 : kCodec == kCodecBUFF ? 5 : 8;
};
template <uint32_t kCodec>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TileWaves<kCodec>::value))) k_decode(
    const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t n_jobs, uint32_t *err) {
  __shared__ __align__(16) double lds[4][kStageStride * kWave];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t stride = gridDim.x * 4;
  ColCursor cur;
  for (uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave); j < n_jobs; j += stride) {
    const ColView cv = cols[cur.at(prefix, n_cols, j)];
    if (cv.codec != kCodec) {
      // wave-uniform: on to this wave's first job past the column (the loop
      // adds the last stride)
      j += (cur.hi - j - 1) / stride * stride;
      continue;
    }
    tile_job_of<kCodec>(cv, j - cur.lo, lds[wave], err);
  }
}

// Persistent decode over the checkpoint-segment jobs of n_cols columns, with
// the same static interleave as k_decode.  A kernel of its own, so that the
// serial bit-stream loops get their own register allocation and LDS (one
// [17][64] output stage per wave).  amdgpu_waves_per_eu(4): the register
// budget at which the kernel needs no scratch; built to fit 5 waves per SIMD
// it spilled, and then decoded Elf wrongly in workgroups that shared a CU
// with an earlier one (the decode kernels are kept spill-free, which a test
// checks).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_decode_seg(
    const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t n_jobs, uint32_t *err) {
  __shared__ __align__(16) double lds[4][(kRun + 1) * kWave];
  __shared__ uint8_t dists[4][kJob128];  // Chimp128 jobs: source distances
  __shared__ double elf_tab[kElfTab];
  if (threadIdx.x < kElfTab) elf_tab_fill(elf_tab, threadIdx.x, c_p10, c_n10);
  __syncthreads();
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t stride = gridDim.x * 4;
  ColCursor cur;
  for (uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave); j < n_jobs; j += stride) {
    const ColView cv = cols[cur.at(prefix, n_cols, j)];
    seg_job_of(cv, j - cur.lo, lds[wave], dists[wave], elf_tab, err);
  }
}

// Both job lists in one persistent launch: job j of the n_seg + n_tile
// combined jobs is segment job floor(j * n_seg / n) when that floor steps
// at j + 1, else tile job j - floor(j * n_seg / n) -- the two kinds spread
// evenly over the combined range, so every wave alternates between the
// HBM-bound tile jobs and the bit-stream loops of the segment jobs.
// Register budget and LDS of k_decode_seg (the tile stage is the segment
// stage's first kStageStride rows).  A candidate of the batch tuner: on a
// batch of 16 tile and 16 XOR columns of 2^28 records it took 31.6-32.3 ms
// at 4-32 blocks per CU against 30.6 ms for k_decode then k_decode_seg (the
// segment jobs share the CU's memory pipeline with the tile jobs rather
// than leaving it idle), so the two kernels stay the untuned default.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_decode_mix(
    const ColView *cols, const uint32_t *tile_prefix, uint32_t n_tile_jobs, const uint32_t *seg_prefix,
    uint32_t n_seg_jobs, uint32_t n_cols, uint32_t *err) {
  static_assert(kStageStride <= kRun + 1, "tile stage inside the segment stage");
  __shared__ __align__(16) double lds[4][(kRun + 1) * kWave];
  __shared__ uint8_t dists[4][kJob128];
  __shared__ double elf_tab[kElfTab];
  if (threadIdx.x < kElfTab) elf_tab_fill(elf_tab, threadIdx.x, c_p10, c_n10);
  __syncthreads();
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t stride = gridDim.x * 4;
  const uint32_t n = n_tile_jobs + n_seg_jobs;
  for (uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave); j < n; j += stride) {
    const uint32_t a = (uint32_t)((uint64_t)j * n_seg_jobs / n);
    const uint32_t b = (uint32_t)((uint64_t)(j + 1) * n_seg_jobs / n);
    if (b > a) {
      const uint32_t ci = find_col(seg_prefix, n_cols, a);
      seg_job_of(cols[ci], a - seg_prefix[ci], lds[wave], dists[wave], elf_tab, err);
    } else
      tile_job(cols, tile_prefix, n_cols, j - a, lds[wave], err);
  }
}

// Query output of segment_wave (the fused queries of k_query_seg): each
// lane keeps, for its own segment, the bitmask bits of its records (all set
// without a bitmask), the records it selects (bitmask and, with `pred`, the
// f64 predicate; bit p = record i0 + p), and their double-double sum (agg 0)
// or min / max order keys (agg 1 / 2, NaN skipped as Rust's f64::min / max
// do).  No record leaves the registers.
struct QueryOut {
  PredF64 pr;
  uint32_t i0, n;   // the lane's first record, the chunk's record count
  bool pred;
  int agg;          // -1 none, 0 sum, 1 min, 2 max
  uint64_t mbits, obits;
  double hi, lo;
  uint64_t kmin, kmax;
  __device__ __forceinline__ void row(uint32_t i, uint32_t, double v, bool act) {
    const uint32_t p = (i - i0) & 63;  // the terminator (i == n) may be record i0 + 64: never selected
    const bool sel = act && i < n && ((mbits >> p) & 1) && (!pred || pred_f64(pr, v));
    obits |= (uint64_t)sel << p;
    if (agg == 0) {
      if (sel) dd_add(hi, lo, v, 0.0);
    } else if (agg > 0 && sel && v == v) {
      kmin = min(kmin, f64_key(v));
      kmax = max(kmax, f64_key(v));
    }
  }
  __device__ __forceinline__ void run_end(uint32_t) {}
};

// Query output of c128_job (the fused queries of a Chimp128 column), flags
// as k_query_seg's: per step one ballot of selection bits (the step's 64
// consecutive records), per record tile (8 steps; a job holds two) one
// double-double partial, the butterfly sum of the lanes' sums; min / max
// keys per lane.
struct C128Query {
  static constexpr bool kNanEnds = false;
  PredF64 pr;
  const uint64_t *mask;
  uint64_t *out_bits;
  double *partials;
  const DevChunk *c;
  bool pred, want_bits;
  int agg;
  double hi, lo;
  uint64_t kmin, kmax;
  uint32_t cnt;  // selected records (wave-uniform)
  __device__ __forceinline__ void put(uint32_t i, uint64_t vb, uint32_t base) {
    const uint32_t n = c->n;
    const uint64_t gi = c->rec_off + i;
    const double v = as_f64(vb);
    const bool m = i < n && (!mask || ((mask[gi >> 6] >> (gi & 63)) & 1));
    const bool sel = m && (!pred || pred_f64(pr, v));
    if (agg == 0) cnt += (uint32_t)__popcll(__ballot(sel));
    if (want_bits) {
      const uint64_t bal = __ballot(sel);
      if (lane_id() == 0 && base < n) {
        // the step's word: the wave's own when aligned and inside the
        // chunk, atomic ORs when it is shared
        const uint64_t g0 = c->rec_off + base;
        unsigned long long *w = reinterpret_cast<unsigned long long *>(out_bits) + (g0 >> 6);
        const uint32_t sh = (uint32_t)(g0 & 63);
        if (sh == 0 && base + kWave <= n) {
          *w = bal;
        } else if (bal) {
          atomicOr(w, (unsigned long long)(bal << sh));
          if (sh) atomicOr(w + 1, (unsigned long long)(bal >> (64 - sh)));
        }
      }
    }
    if (agg == 0) {
      if (sel) dd_add(hi, lo, v, 0.0);
      // the last step with records of its tile: the tile's partial
      if (base < n && ((base % kTile) == kTile - kWave || base + kWave >= n)) {
        const uint32_t lane = lane_id();
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
          const double oh = __shfl_xor(hi, d, kWave), ol = __shfl_xor(lo, d, kWave);
          if ((lane & d) == 0) {
            dd_add(hi, lo, oh, ol);
          } else {
            double h2 = oh, l2 = ol;
            dd_add(h2, l2, hi, lo);
            hi = h2;
            lo = l2;
          }
        }
        if (lane == 0) {
          double *pp = partials + 2ull * (c->rtile0 + base / kTile);
          pp[0] = hi;
          pp[1] = lo;
        }
        hi = lo = 0.0;
      }
    } else if (agg > 0 && sel && v == v) {
      kmin = min(kmin, f64_key(v));
      kmax = max(kmax, f64_key(v));
    }
  }
};

// Wave min / max of the lanes' keys into chunk accumulators acc[2] / acc[3].
__device__ __forceinline__ void keys_to_acc(int agg, uint64_t kmin, uint64_t kmax, uint64_t *acc) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    kmin = min(kmin, (uint64_t)__shfl_xor((unsigned long long)kmin, d, kWave));
    kmax = max(kmax, (uint64_t)__shfl_xor((unsigned long long)kmax, d, kWave));
  }
  if (lane_id() == 0) {
    if (agg == 1) atomicMin(reinterpret_cast<unsigned long long *>(acc) + 2, (unsigned long long)kmin);
    else atomicMax(reinterpret_cast<unsigned long long *>(acc) + 3, (unsigned long long)kmax);
  }
}

// Fused queries over the checkpoint-segment jobs of every plan of a group
// with kPfSegs whose column has codec kCodec (Gorilla / Chimp / Elf /
// ElfOnChimp / Chimp128; the plans' segment jobs concatenated, prefix over
// all plans of the group, one launch per codec present: a wave that reaches
// a plan of another codec steps past its jobs at once): the records are
// decoded as in k_decode_seg and consumed in registers.  Per plan: kPfPred
// applies the f64 predicate, kPfBits writes the selected records' bits to
// a.out_bits (zeroed before), the aggregate (flag bits 4-5) goes to the
// per-record-tile double-double partials (sum; each the butterfly sum of its
// 8 segments' lane sums) or the chunk accumulators (min / max), and with
// kPfAggSel the selected records are counted per chunk (acc[1]).  Chunks
// with a.chunk_sel 0 are skipped.  One kernel per codec, so that each
// keeps its own registers: Gorilla and Chimp records never touch LDS and run
// at the waves per SIMD their registers allow; Chimp128 jobs need the
// [kSeg128][64] stage and the distances (four blocks per CU by LDS).
template <uint32_t kCodec>
struct QuerySegWaves {
  static constexpr int value = kCodec == kCodecGorilla || kCodec == kCodecChimp ? 5 : 4;
};
template <uint32_t kCodec>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(QuerySegWaves<kCodec>::value))) k_query_seg(
    const PlanDesc *__restrict__ plans, const uint32_t *__restrict__ prefix, uint32_t n_plans, uint32_t n_jobs,
    uint32_t *err) {
  constexpr bool kC128 = kCodec == kCodecChimp128;
  constexpr bool kTab = kCodec == kCodecElf || kCodec == kCodecElfOnChimp;
  __shared__ __align__(16) double lds[4][kC128 ? kSeg128 * kWave : 1];  // Chimp128 jobs: the payload / value stage
  __shared__ uint8_t dists[4][kC128 ? kJob128 : 1];                    // and the source distances
  __shared__ double elf_tab[kTab ? kElfTab : 1];
  if (kTab) {
    if (threadIdx.x < kElfTab) elf_tab_fill(elf_tab, threadIdx.x, c_p10, c_n10);
    __syncthreads();
  }
  const uint32_t lane = lane_id();
  const uint32_t stride = gridDim.x * 4;
  ColCursor cur;  // over the plans
  for (uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / kWave); j < n_jobs; j += stride) {
    const uint32_t pi = cur.at(prefix, n_plans, j);
    const PlanDesc &P = plans[pi];
    if (P.scheme != kCodec) {
      // wave-uniform: on to this wave's first job past the plan (the loop
      // adds the last stride)
      j += (cur.hi - j - 1) / stride * stride;
      continue;
    }
    const PredF64 pr = pred_of(P.a);
    const uint64_t *mask = P.a.mask;
    uint64_t *out_bits = P.a.out_bits;
    double *partials = P.a.partials;
    uint64_t *acc = P.a.acc;
    const uint32_t flags = P.flags;
    const bool want_bits = (flags & kPfBits) != 0;
    const bool pred = (flags & kPfPred) != 0;
    const int agg = (int)((flags >> kPfAggShift) & 3u) - 1;  // -1 none, 0 sum, 1 min, 2 max
    const bool count = agg == 0 && (flags & kPfAggSel) != 0;
    const ColView cv = *static_cast<const ColView *__restrict__>(P.view);
    const Job jb = cv.seg_jobs[j - cur.lo];
    if (!P.a.chunk_sel[jb.chunk]) continue;
    const DevChunk c = cv.chunks[jb.chunk];
    if constexpr (kC128) {
      const uint32_t wave = threadIdx.x / kWave;
      C128Query q{pr, mask, out_bits, partials, &c, pred, want_bits, agg, 0.0, 0.0, ~0ull, 0, 0};
      const uint32_t e = c128_job<false>(cv, c, jb.tile, jb.col, lds[wave], dists[wave], q);
      if (e) atomicOr(err, e);
      if (agg > 0) keys_to_acc(agg, q.kmin, q.kmax, acc + 4ull * jb.chunk);
      if (count && lane == 0 && q.cnt) atomicAdd((unsigned long long *)(acc + 4ull * jb.chunk) + 1, (unsigned long long)q.cnt);
    } else {
      const uint32_t i0 = (jb.tile + lane) * kSeg;
      const uint64_t g = c.rec_off + i0;
      uint64_t m = ~0ull;
      if (mask && i0 < c.n) {  // the 64 bitmask bits from record g (the bitmask has a spare word)
        const uint32_t sh = (uint32_t)(g & 63);
        m = mask[g >> 6] >> sh;
        if (sh) m |= mask[(g >> 6) + 1] << (64 - sh);
      }
      QueryOut q{pr, i0, c.n, pred, agg, m, 0, 0.0, 0.0, ~0ull, 0};
      const uint32_t e = c.codec == kCodec ? segment_wave<kCodec>(cv, c, jb.tile, jb.col, elf_tab, q) : kErrCorrupt;
      if (e) atomicOr(err, e);
      if (want_bits && i0 < c.n) {
        // a word wholly inside this segment is the lane's own; one shared
        // with another segment or chunk takes atomic ORs
        unsigned long long *w = reinterpret_cast<unsigned long long *>(out_bits) + (g >> 6);
        const uint32_t sh = (uint32_t)(g & 63);
        if (sh == 0 && i0 + kSeg <= c.n) {
          *w = q.obits;
        } else if (q.obits) {
          atomicOr(w, (unsigned long long)(q.obits << sh));
          if (sh) atomicOr(w + 1, (unsigned long long)(q.obits >> (64 - sh)));
        }
      }
      if (agg == 0) {
        // the 8 segments of a record tile: butterfly in a lane-independent order
#pragma unroll
        for (int d = 1; d < 8; d <<= 1) {
          const double oh = __shfl_xor(q.hi, d, kWave), ol = __shfl_xor(q.lo, d, kWave);
          if ((lane & d) == 0) {
            dd_add(q.hi, q.lo, oh, ol);
          } else {
            double h2 = oh, l2 = ol;
            dd_add(h2, l2, q.hi, q.lo);
            q.hi = h2;
            q.lo = l2;
          }
        }
        if ((lane & 7) == 0 && i0 < c.n) {
          double *pp = partials + 2ull * (c.rtile0 + (jb.tile + lane) / 8);
          pp[0] = q.hi;
          pp[1] = q.lo;
        }
        if (count) {
          const uint64_t n = wave_sum_u64(i0 < c.n ? (uint64_t)__popcll(q.obits) : 0ull);
          if (lane == 0 && n) atomicAdd((unsigned long long *)(acc + 4ull * jb.chunk) + 1, (unsigned long long)n);
        }
      } else if (agg > 0) {
        keys_to_acc(agg, q.kmin, q.kmax, acc + 4ull * jb.chunk);
      }
    }
  }
}

// kWrite: decode to cv.out; kIndex: write checkpoints to c.xidx.  Block b
// runs serial job b of the launch-wide list (prefix over n_serial_jobs).
template <bool kWrite, bool kIndex>
__global__ void __launch_bounds__(64) k_serial(const ColView *cols, const uint32_t *prefix, uint32_t n_cols,
                                              const Job *single, uint32_t *err) {
  Job jb;
  ColView cv;
  if (single) {  // one explicit job on column 0 (per-chunk comparator)
    jb = *single;
    cv = cols[0];
  } else {
    const uint32_t ci = find_col(prefix, n_cols, blockIdx.x);
    cv = cols[ci];
    jb = cv.serial_jobs[blockIdx.x - prefix[ci]];
  }
  const uint32_t lane = threadIdx.x;
  if (lane >= jb.tile) return;
  const uint32_t ci = jb.chunk + lane;
  const DevChunk c = cv.chunks[ci];
  if (!kIndex && cv.chunk_err && cv.chunk_err[ci]) return;
  const uint8_t *data = reinterpret_cast<const uint8_t *>(c.data);
  const DirectSink out{cv.out + c.rec_off};
  XorCkpt *ck = kIndex ? reinterpret_cast<XorCkpt *>(c.xidx) : nullptr;
  const uint32_t bits = c.size * 8u;
  uint32_t e;
  if (c.codec == kCodecChimp128) {
    e = kIndex ? chimp128_index(data, bits, c.n, ck) : kErrCorrupt;
  } else {
    SerialState s;
    xor_state_start(s, c.codec, data);
    e = xor_run<kWrite>(c.codec, s, bits, c.n, 0, c.n, out, ck);
    if (e == kNanEnd) {
      // the chunk ends at its first NaN: the host shortens it to s.end records
      if (cv.chunk_end) cv.chunk_end[ci] = s.end;
      e = kErrNone;
    }
  }
  if (e) {
    atomicOr(err, e);
    if (cv.chunk_err) atomicOr(cv.chunk_err + ci, e);
  }
}

template <bool kWrite>
__global__ void __launch_bounds__(64) k_chimp128(const ColView *cols, const uint32_t *prefix, uint32_t n_cols,
                                                const Job *single, uint32_t *err) {
  __shared__ uint64_t ring[128 * kWave];
  Job jb;
  ColView cv;
  if (single) {
    jb = *single;
    cv = cols[0];
  } else {
    const uint32_t ci = find_col(prefix, n_cols, blockIdx.x);
    cv = cols[ci];
    jb = cv.serial_jobs[blockIdx.x - prefix[ci]];
  }
  const uint32_t lane = threadIdx.x;
  if (lane >= jb.tile) return;
  const uint32_t ci = jb.chunk + lane;
  const DevChunk c = cv.chunks[ci];
  if (kWrite && cv.chunk_err && cv.chunk_err[ci]) return;
  uint32_t end = 0;
  uint32_t e = chimp128_run<kWrite>(reinterpret_cast<const uint8_t *>(c.data), c.size * 8u, c.n,
                                    cv.out + c.rec_off, ring + lane, &end);
  if (e == kNanEnd) {  // the chunk ends at its first NaN (see k_serial)
    if (cv.chunk_end) cv.chunk_end[ci] = end;
    e = kErrNone;
  }
  if (e) {
    atomicOr(err, e);
    if (cv.chunk_err) atomicOr(cv.chunk_err + ci, e);
  }
}

// Load-time ring checkpoints of the Chimp128 column cols[0], decoded into
// cols[0].out (one wave per segment job): the values of the 128 records
// before the job's first record, zeros before record 0 (and for chunks the
// index pass rejected).
__global__ void __launch_bounds__(256) k_c128_ring(const ColView *cols, uint32_t n_jobs, uint64_t *ring) {
  const uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / kWave);
  if (j >= n_jobs) return;
  const ColView cv = cols[0];
  const Job jb = cv.seg_jobs[j];
  const DevChunk c = cv.chunks[jb.chunk];
  const bool ok = !(cv.chunk_err && cv.chunk_err[jb.chunk]);
  const int64_t j0 = (int64_t)jb.tile * kSeg128;
  for (uint32_t u = lane_id(); u < 128; u += kWave) {
    const int64_t r = j0 - 128 + u;
    ring[128ull * jb.col + u] = ok && r >= 0 ? as_u64(cv.out[c.rec_off + r]) : 0;
  }
}

// Output of c128_job for the load-time ring build: the values of the job's
// last 128 records become the ring checkpoint of the chunk's next job, and
// the first NaN before the chunk's footer count is where the reference ends
// the chunk (chimp_n.rs:219-266), so it is kept instead of flagged.
struct C128Ring {
  static constexpr bool kNanEnds = true;
  uint64_t *next;  // ring of the chunk's next job, null for its last job
  uint32_t tail0;  // first record held by that ring
  uint32_t n;
  uint32_t nan;    // smallest record < n whose value is NaN (this lane's)
  __device__ __forceinline__ void put(uint32_t i, uint64_t v, uint32_t) {
    if (next && i - tail0 < 128u) next[i - tail0] = v;
    if (i < n && v == kNan) nan = min(nan, i);
  }
};

// One step of the load-time Chimp128 ring build over the column cols[0], a
// wavefront over job indexes: wave w decodes job `step` of chunk w (its ring
// written by the previous step; job 0 reads none), writes the next job's
// ring and lowers the chunk's end to its first NaN.  job_first: per chunk its
// first segment job, nc + 1 entries.  Chunks the index pass rejected keep
// zero rings and are never decoded.
__global__ void __launch_bounds__(256) k_c128_ring_step(const ColView *cols, const uint32_t *job_first, uint32_t nc,
                                                        uint32_t step, uint64_t *ring, uint32_t *err) {
  __shared__ __align__(16) double lds[4][(kRun + 1) * kWave];
  __shared__ uint8_t dists[4][kJob128];
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t ci = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  if (ci >= nc) return;
  const uint32_t j = job_first[ci] + step;
  if (j >= job_first[ci + 1]) return;  // wave-uniform
  const ColView cv = cols[0];
  if (cv.chunk_err && cv.chunk_err[ci]) return;
  const Job jb = cv.seg_jobs[j];
  const DevChunk c = cv.chunks[ci];
  C128Ring o{j + 1 < job_first[ci + 1] ? ring + 128ull * (jb.col + 1) : nullptr,
             jb.tile * kSeg128 + kJob128 - 128, c.n, ~0u};
  uint32_t e = c128_job<true>(cv, c, jb.tile, jb.col, lds[wave], dists[wave], o);
  uint32_t nan = o.nan;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) nan = min(nan, (uint32_t)__shfl_xor((int)nan, d, kWave));
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) e |= (uint32_t)__shfl_xor((int)e, d, kWave);
  if (lane_id() == 0) {
    if (nan != ~0u && cv.chunk_end) atomicMin(cv.chunk_end + ci, nan);
    if (e) {
      atomicOr(err, e);
      if (cv.chunk_err) atomicOr(cv.chunk_err + ci, e);
    }
  }
}

// Tile sums for the load-time index: DeltaSprintz sum of deltas, RLE sum of runs.
__global__ void __launch_bounds__(256) k_tile_sums(const ColView *cols, const Job *jobs, uint32_t n_jobs) {
  const uint32_t j = blockIdx.x * 4 + threadIdx.x / kWave;
  if (j >= n_jobs) return;
  const Job jb = jobs[j];
  const ColView cv = cols[jb.col];
  const DevChunk c = cv.chunks[jb.chunk];
  int64_t s = 0;
  if (c.codec == kCodecSprintz) s = sprintz_tile_sum(c, jb.tile);
  else if (c.codec == kCodecRLE) s = (int64_t)rle_tile_sum(c, jb.tile);
  else return;
  if (lane_id() == 0) cv.tile_index[c.tile0 + jb.tile] = s;
}

// One thread per chunk: exclusive scan of its tile sums in place.
__global__ void k_tile_scan(const ColView *cols, uint32_t col, uint32_t n_chunks, uint32_t *err) {
  const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
  if (ci >= n_chunks) return;
  const ColView cv = cols[col];
  const DevChunk c = cv.chunks[ci];
  int64_t *idx = cv.tile_index + c.tile0;
  if (c.codec == kCodecSprintz) {
    const uint32_t nt = (c.n + kTile - 1) / kTile;
    uint64_t acc = (uint64_t)c.base;
    for (uint32_t t = 0; t < nt; ++t) {
      const uint64_t s = (uint64_t)idx[t];
      idx[t] = (int64_t)acc;
      acc += s;
    }
  } else if (c.codec == kCodecRLE) {
    const uint32_t nt = (uint32_t)((c.fields + kTile - 1) / kTile);
    uint64_t acc = 0;
    for (uint32_t t = 0; t < nt; ++t) {
      const uint64_t s = (uint64_t)idx[t];
      idx[t] = (int64_t)acc;
      acc += s;
    }
    if (acc != c.n) {  // the run total must match the footer record count
      atomicOr(err, (uint32_t)kErrCount);
      if (cv.chunk_err) atomicOr(cv.chunk_err + ci, (uint32_t)kErrCount);
    }
  }
}

// Lane l of the wave for segment job jb of view cv (Gorilla, Chimp, Elf):
// the stream of segment jb.tile + l as an ilv lane -- its first chunk block
// (first_blk) and its block count (0 for a lane past the chunk's segments or
// a chunk the index pass rejected).  The segment's bits run from its
// checkpoint to the next (the last segment's to the chunk's end), clamped to
// the chunk.
__device__ __forceinline__ uint32_t ilv_lane(const ColView &cv, const Job &jb, uint32_t &first_blk) {
  const DevChunk c = cv.chunks[jb.chunk];
  const uint32_t n_segs = (c.n + kSeg - 1) / kSeg;
  const uint32_t seg = jb.tile + lane_id();
  const uint32_t bits = c.size * 8u;
  first_blk = 0;
  if (seg >= n_segs || (cv.chunk_err && cv.chunk_err[jb.chunk])) return 0;
  const XorCkpt *ck = reinterpret_cast<const XorCkpt *>(c.xidx);
  const uint32_t a = min(ck[seg].bitpos, bits);
  const uint32_t b = seg + 1 < n_segs ? min(max(ck[seg + 1].bitpos, a), bits) : bits;
  first_blk = a >> 7;
  return ilv_lane_blocks(a, b);
}

// Rows of each segment job of cols[0] (one wave per job): its longest lane.
__global__ void __launch_bounds__(256) k_ilv_rows(const ColView *cols, uint32_t n_jobs, uint32_t *rows) {
  const uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / kWave);
  if (j >= n_jobs) return;
  const ColView cv = cols[0];
  const Job jb = cv.seg_jobs[j];
  uint32_t fb;
  uint32_t m = ilv_lane(cv, jb, fb);
  for (int o = kWave / 2; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if (lane_id() == 0) rows[j] = m;
}

// Fills the rows of each segment job of cols[0] (jb.col = its first row,
// rows[j] of them): block r of lane l is the lane's chunk block first + r
// while r is below its block count, zero after.  Each store covers a whole
// 1 KiB row.
__global__ void __launch_bounds__(256) k_ilv_copy(const ColView *cols, uint32_t n_jobs, const uint32_t *rows,
                                                  uint4 *ilv) {
  const uint32_t j = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + threadIdx.x / kWave);
  if (j >= n_jobs) return;
  const ColView cv = cols[0];
  const Job jb = cv.seg_jobs[j];
  uint32_t fb;
  const uint32_t nb = ilv_lane(cv, jb, fb);
  const EBI_GLOBAL uint4 *src = as_global(reinterpret_cast<const uint4 *>(cv.chunks[jb.chunk].data)) + fb;
  uint4 *dst = ilv + (size_t)jb.col * kIlvRowBlocks + lane_id();
  const uint32_t n_rows = rows[j];
  for (uint32_t r = 0; r < n_rows; ++r) {
    uint4 x = make_uint4(0, 0, 0, 0);
    if (r < nb) x = src[r];
    dst[(size_t)r * kIlvRowBlocks] = x;
  }
}

__global__ void k_zero_u32(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------- launchers
namespace launch {

hipError_t set_pow10(const double *p10, const double *n10) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_p10), p10, sizeof(double) * 400);
  if (e != hipSuccess) return e;
  return hipMemcpyToSymbol(HIP_SYMBOL(c_n10), n10, sizeof(double) * 400);
}

void zero_u32(hipStream_t s, uint32_t *p, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

void decode(hipStream_t s, const ColView *cols, uint32_t n_cols, const uint32_t *tile_prefix, uint32_t n_tile_jobs,
            const uint32_t *seg_prefix, uint32_t n_seg_jobs, uint32_t *err, uint32_t max_blocks, hipEvent_t mid,
            bool mixed, uint32_t codecs, hipStream_t side, hipEvent_t fork, hipEvent_t join,
            uint32_t side_blocks) {
  auto grid = [&](uint32_t jobs) {
    const uint32_t blocks = (jobs + 3) / 4;
    return max_blocks && blocks > max_blocks ? max_blocks : blocks;
  };
  // the combined job index and its grid stride stay below 2^32: n < 2^31 - 2
  if (mixed && n_tile_jobs && n_seg_jobs && (uint64_t)n_tile_jobs + n_seg_jobs < (1ull << 31) - 2) {
    hipLaunchKernelGGL(k_decode_mix, dim3(grid(n_tile_jobs + n_seg_jobs)), dim3(256), 0, s, cols, tile_prefix,
                       n_tile_jobs, seg_prefix, n_seg_jobs, n_cols, err);
    if (mid) hipEventRecord(mid, s);
    return;
  }
  if (!codecs) codecs = kTileCodecs;
  const bool beside = side && fork && join && side_blocks && n_tile_jobs && n_seg_jobs &&
                      (codecs & (1u << kCodecPlain));
  if (beside) codecs &= ~(1u << kCodecPlain);
  if (n_tile_jobs) {
    const dim3 g(grid(n_tile_jobs));
    if (codecs & (1u << kCodecPlain))
      hipLaunchKernelGGL(k_decode<kCodecPlain>, g, dim3(256), 0, s, cols, tile_prefix, n_cols, n_tile_jobs, err);
    if (codecs & (1u << kCodecBUFF))
      hipLaunchKernelGGL(k_decode<kCodecBUFF>, g, dim3(256), 0, s, cols, tile_prefix, n_cols, n_tile_jobs, err);
    if (codecs & (1u << kCodecRLE))
      hipLaunchKernelGGL(k_decode<kCodecRLE>, g, dim3(256), 0, s, cols, tile_prefix, n_cols, n_tile_jobs, err);
    if (codecs & (1u << kCodecSprintz))
      hipLaunchKernelGGL(k_decode<kCodecSprintz>, g, dim3(256), 0, s, cols, tile_prefix, n_cols, n_tile_jobs, err);
  }
  if (mid) hipEventRecord(mid, s);
  if (beside) {
    hipEventRecord(fork, s);
    hipStreamWaitEvent(side, fork, 0);
    const uint32_t blocks = (n_tile_jobs + 3) / 4;
    hipLaunchKernelGGL(k_decode<kCodecPlain>, dim3(blocks < side_blocks ? blocks : side_blocks), dim3(256), 0, side,
                       cols, tile_prefix, n_cols, n_tile_jobs, err);
  }
  if (n_seg_jobs)
    hipLaunchKernelGGL(k_decode_seg, dim3(grid(n_seg_jobs)), dim3(256), 0, s, cols, seg_prefix, n_cols,
                       n_seg_jobs, err);
  if (beside) {
    hipEventRecord(join, side);
    hipStreamWaitEvent(s, join, 0);
  }
}

void query_seg(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_jobs,
               uint32_t *err, uint32_t max_blocks, uint32_t codecs) {
  if (!n_jobs) return;
  uint32_t blocks = (n_jobs + 3) / 4;
  if (max_blocks && blocks > max_blocks) blocks = max_blocks;
  if (!codecs) codecs = kSegCodecs;
  const dim3 g(blocks);
  if (codecs & (1u << kCodecGorilla))
    hipLaunchKernelGGL(k_query_seg<kCodecGorilla>, g, dim3(256), 0, s, plans, prefix, n_plans, n_jobs, err);
  if (codecs & (1u << kCodecChimp))
    hipLaunchKernelGGL(k_query_seg<kCodecChimp>, g, dim3(256), 0, s, plans, prefix, n_plans, n_jobs, err);
  if (codecs & (1u << kCodecElf))
    hipLaunchKernelGGL(k_query_seg<kCodecElf>, g, dim3(256), 0, s, plans, prefix, n_plans, n_jobs, err);
  if (codecs & (1u << kCodecElfOnChimp))
    hipLaunchKernelGGL(k_query_seg<kCodecElfOnChimp>, g, dim3(256), 0, s, plans, prefix, n_plans, n_jobs, err);
  if (codecs & (1u << kCodecChimp128))
    hipLaunchKernelGGL(k_query_seg<kCodecChimp128>, g, dim3(256), 0, s, plans, prefix, n_plans, n_jobs, err);
}

void serial(hipStream_t s, const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t n_jobs,
            const Job *single, uint32_t *err, bool write, bool index) {
  if (!n_jobs) return;
  if (write && index)
    hipLaunchKernelGGL((k_serial<true, true>), dim3(n_jobs), dim3(64), 0, s, cols, prefix, n_cols, single, err);
  else if (write)
    hipLaunchKernelGGL((k_serial<true, false>), dim3(n_jobs), dim3(64), 0, s, cols, prefix, n_cols, single, err);
  else
    hipLaunchKernelGGL((k_serial<false, true>), dim3(n_jobs), dim3(64), 0, s, cols, prefix, n_cols, single, err);
}

void c128_ring_step(hipStream_t s, const ColView *cols, const uint32_t *job_first, uint32_t nc, uint32_t step,
                    uint64_t *ring, uint32_t *err) {
  if (nc) hipLaunchKernelGGL(k_c128_ring_step, dim3((nc + 3) / 4), dim3(256), 0, s, cols, job_first, nc, step, ring, err);
}
void c128_ring(hipStream_t s, const ColView *cols, uint32_t n_jobs, uint64_t *ring) {
  if (n_jobs) hipLaunchKernelGGL(k_c128_ring, dim3((n_jobs + 3) / 4), dim3(256), 0, s, cols, n_jobs, ring);
}

void chimp128(hipStream_t s, const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t n_jobs,
              const Job *single, uint32_t *err, bool write) {
  if (!n_jobs) return;
  if (write)
    hipLaunchKernelGGL((k_chimp128<true>), dim3(n_jobs), dim3(64), 0, s, cols, prefix, n_cols, single, err);
  else
    hipLaunchKernelGGL((k_chimp128<false>), dim3(n_jobs), dim3(64), 0, s, cols, prefix, n_cols, single, err);
}

void ilv_rows(hipStream_t s, const ColView *cols, uint32_t n_jobs, uint32_t *rows) {
  if (n_jobs) hipLaunchKernelGGL(k_ilv_rows, dim3((n_jobs + 3) / 4), dim3(256), 0, s, cols, n_jobs, rows);
}

void ilv_copy(hipStream_t s, const ColView *cols, uint32_t n_jobs, const uint32_t *rows, uint4 *ilv) {
  if (n_jobs) hipLaunchKernelGGL(k_ilv_copy, dim3((n_jobs + 3) / 4), dim3(256), 0, s, cols, n_jobs, rows, ilv);
}

void tile_index(hipStream_t s, const ColView *cols, uint32_t col, const Job *jobs, uint32_t n_jobs,
                uint32_t n_chunks, uint32_t *err) {
  if (n_jobs) hipLaunchKernelGGL(k_tile_sums, dim3((n_jobs + 3) / 4), dim3(256), 0, s, cols, jobs, n_jobs);
  if (n_chunks)
    hipLaunchKernelGGL(k_tile_scan, dim3((n_chunks + 255) / 256), dim3(256), 0, s, cols, col, n_chunks, err);
}

}  // namespace launch
}  // namespace ebi
