// Resident column layout and work descriptors shared by the host loader and
// the gfx950 kernels of libebi_hip.
//
// HBM layout of a resident column (one allocation per column):
//   * every chunk body starts on a 256-byte boundary and is followed by at
//     least kPad zero bytes, so wave loads never cross into another chunk
//     and bit readers may over-read a little past the end;
//   * BUFF bodies are re-laid out: each byte plane and the LSB tail start on
//     their own 256-byte boundary (plane_stride = round_up(n, 256)), the
//     20-byte header lives in the descriptor;
//   * DeltaSprintz bodies start at the first W-bit field (byte 13 of the
//     chunk), the 13-byte header lives in the descriptor;
//   * RLE bodies are split into a u8 run array and an 8-byte aligned f64
//     value array (structure of arrays);
//   * Uncompressed, Gorilla, Chimp, Chimp128 and Elf bodies are the chunk
//     bytes unchanged.
// Next to the bodies the column keeps a small tile index computed once at
// load time by the GPU (prefix state every kTile records or fields), and
// XOR columns a second copy of their streams, lane-interleaved for the
// segment decode (ilv, below; its own allocation).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ebi {

constexpr uint32_t kAlign = 256;
constexpr uint32_t kPad = 256;
constexpr uint32_t kTile = 512;  // records (fields for RLE) per wave job: 64 lanes x 8
constexpr uint32_t kSeg = 64;    // records per checkpoint of the XOR index
constexpr uint32_t kSeg128 = 16; // records per checkpoint of the Chimp128 index
constexpr uint32_t kJob128 = kSeg128 * 64;  // records of a Chimp128 segment job (64 segments)
constexpr uint32_t kTilesPerJob = 4;  // consecutive tiles of one chunk per decode job
constexpr uint32_t kSpanTiles = 16;   // most consecutive tiles of one chunk per query span job
constexpr uint32_t kSpanMinWaves = 8192;  // span jobs a column's query aims for at least
constexpr uint32_t kWaveLanes = 64;  // gfx950 wavefront size

enum Codec : uint32_t {
  kCodecPlain = 0,
  kCodecRLE = 1,
  kCodecBUFF = 2,
  kCodecGorilla = 3,
  kCodecChimp = 4,
  kCodecChimp128 = 5,
  kCodecElfOnChimp = 6,
  kCodecElf = 7,
  kCodecSprintz = 8,
};

struct alignas(16) DevChunk {
  uint64_t data;      // absolute device address of the chunk body
  uint64_t rec_off;   // global id of the first record (ChunkFooter::logical_offset)
  uint32_t n;         // records in the chunk (footer geometry)
  uint32_t size;      // bytes of the chunk in the file
  uint32_t codec;     // Codec
  uint32_t tile0;     // first tile of this chunk in the column's tile index
  // BUFF: B (bits), r (tail bits), n_planes, frac, base, plane_stride
  // Sprintz: W (bits), scale, q0
  // RLE: fields, value array offset from data
  uint32_t bits;      // BUFF B / Sprintz W
  uint32_t r;         // BUFF tail bits
  uint32_t n_planes;  // BUFF byte planes
  uint32_t frac;      // BUFF fractional bits
  int64_t base;       // BUFF base fixed value / Sprintz q0
  uint32_t scale;     // Sprintz scale (u32 from the header)
  uint32_t plane_stride;  // BUFF bytes between planes
  uint64_t fields;    // RLE field count
  uint64_t values_off;    // RLE: byte offset of the f64 array from data
  uint64_t xidx;      // absolute address of this chunk's XOR checkpoints (0 = none)
  uint32_t rtile0;    // first record tile (kTile records) of this chunk in the column
  uint32_t pad0;
  uint64_t pad1;
};
static_assert(sizeof(DevChunk) == 112, "DevChunk layout");

// Per-chunk query parameters computed on the host from the chunk header and
// the predicate (BUFF fixed-point / DeltaSprintz quantized domain).
enum SideMode : uint32_t {
  kSideNone = 0,   // RangeValue::None (no bound on this side)
  kSideAll = 1,    // early exit: every record of the chunk, bitmask ignored
  kSideEmpty = 2,  // early exit: no record
  kSideCmp = 3,    // compare per record (restricted to the bitmask)
};
struct alignas(16) ChunkPred {
  uint32_t lmode, rmode;   // BUFF range sides / Eq (lmode) ; Sprintz: unused
  uint32_t lincl, rincl;   // inclusive flags
  uint64_t ldp, rdp;       // BUFF: predicate delta (fixed - base); Sprintz: quantized operands
  uint32_t eq_u8;          // BUFF Eq with B < 8 and no bitmask: compare with (u8)dp
  uint32_t pad;
  uint64_t pad1;
};
static_assert(sizeof(ChunkPred) == 48, "ChunkPred layout");

// Per-call query arguments (passed by value).
struct QueryArgs {
  const uint64_t *mask;      // input bitset over global ids (nullable)
  uint64_t *out_bits;        // output bitset (filter / select)
  const double *values;      // decoded records, global index (default codecs)
  const uint32_t *chunk_sel; // per chunk: 1 if selected (chunk_id, pruning)
  const ChunkPred *cpred;    // per chunk predicate parameters (BUFF, Sprintz)
  uint64_t *acc;             // per chunk: sum, count, min key, max key
  double *partials;          // per record tile: (hi, lo) double-double sum
  int32_t kind, start_kind, end_kind, op;  // predicate (f64 domain) and aggregate op
  double value, start, end;
};

// XOR-codec checkpoint (every kSeg records): enough state to restart the
// serial decoder at record k*kSeg.
struct alignas(16) XorCkpt {
  uint64_t value;   // bit pattern of the previous record (prefix-XOR state)
  uint32_t bitpos;  // bit offset of the record's first bit within the chunk
  uint8_t a, b;     // Gorilla L, T | Chimp stored lead | Elf L, T
  int8_t beta;      // Elf last beta* (-1 = unset)
  uint8_t pad;
};
static_assert(sizeof(XorCkpt) == 16, "XorCkpt layout");

// One wave's unit of work in a decode launch.
enum JobKind : uint32_t {
  kJobTile = 0,    // tile(s) of a tile codec (plain, BUFF, Sprintz, RLE); a decode
                   // job's `kind` field holds its tile count instead
  kJobSerial = 1,  // up to 64 XOR chunks, one lane per chunk, serial decode
  kJobSeg = 2,     // 64 checkpoint segments of one XOR chunk, one lane each
};

struct alignas(16) Job {
  uint32_t kind;
  uint32_t col;    // column slot in the launch (output pointer / table);
                   // kJobSeg: first row of the job in the column's ilv
                   // (Chimp128: the job's index in its column, for c128_ring)
  uint32_t chunk;  // chunk index within the column (first chunk for kJobSerial)
  uint32_t tile;   // tile / first segment index; lane count for kJobSerial
};

// Lane-interleaved segment streams (`ilv`) of an XOR column: the 64
// segments of a segment job are read by the 64 lanes of a wave at once,
// each sequentially in 16-byte blocks, so the job's streams are stored
// interleaved: row r of the job (1 KiB) holds block r of every lane's
// stream, lane l's at block r * 64 + l.  Lane l's stream starts with the
// chunk block holding its segment's first bit and runs to two blocks past
// the block holding its last bit (the window holds two blocks and requests
// the next); the job has as many rows as its longest lane stream (shorter
// ones are zero-filled).  A wave's loads then cover whole 128-byte lines of
// few rows instead of one line per lane.
constexpr uint32_t kIlvRowBlocks = 64;
__host__ __device__ inline uint32_t ilv_lane_blocks(uint32_t first_bit, uint32_t end_bit) {
  return (end_bit >> 7) + 3 - (first_bit >> 7);
}


// Per-launch view of one column.  A launch over several columns walks the
// concatenation of their job lists; `prefix` arrays (one entry per column
// plus one) map a launch-wide job id to (column, local job).
struct alignas(16) ColView {
  const DevChunk *chunks;
  int64_t *tile_index;     // per-tile prefix state (Sprintz q, RLE start record)
  double *out;             // decode output base (record r -> out[r])
  uint32_t *chunk_err;     // per-chunk DevError flags found by the index pass
  const Job *tile_jobs;    // resident: kTile-record jobs of tile codecs
  const Job *seg_jobs;     // resident: checkpoint-segment jobs of XOR codecs
  const Job *serial_jobs;  // resident: lane-per-chunk jobs (Chimp128, cold XOR)
  uint64_t n_records;
  uint32_t n_chunks;
  uint32_t n_tile_jobs;
  uint32_t n_seg_jobs;
  uint32_t n_serial_jobs;
  // Chimp128 (null for other codecs): per segment job, the values of the
  // 128 records before its first record (the ring a job starts from; zeros
  // before record 0), at c128_ring[128 * Job::col]
  const uint64_t *c128_ring;
  // per chunk: records before the chunk's first NaN end marker, written by
  // the load-time passes only when that is below the footer's record count
  // (the reference stops a chunk at its first NaN, gorilla.rs:149-152)
  uint32_t *chunk_end;
  // BUFF decode precision of this view (Decoder::with_precision,
  // api/decoder.rs:159-161): -1 full, else 0..12 (buff_tile)
  int32_t precision;
  uint32_t codec;    // the column's scheme (the codec of every chunk)
  const uint4 *ilv;  // lane-interleaved segment streams (Gorilla, Chimp, Elf; null before they are built)
};
static_assert(sizeof(ColView) == 112, "ColView layout");

// Parallel XOR index (kernels/xindex.hpp): a chunk's stream is cut into
// spans of kSpanBits bits (span_count of them: the last one also owns a
// record that would start at the stream's end); one candidate walk covers
// kSpansPerLane spans; Gorilla's candidate walks start in the state the
// chunk's walker reaches at span kHintSpans.
constexpr uint32_t kSpanBits = 4096;
constexpr uint32_t kSpansPerLane = 4;
constexpr uint32_t kHintSpans = 4;
__host__ __device__ inline uint32_t span_count(uint32_t bits) { return bits / kSpanBits + 1; }
// Per span of a chunk's stream:
// the result of parsing it exactly from an entry (the next span's entry,
// the XOR of its records' payloads, its record count, its last beta* or
// -1, the error that ended the exact walk or 0) and its true entry with the
// record index, value and beta* before it; per chunk the walker's state
// (true entry, value, record index and beta* at its frontier span).
struct SpanSum {
  uint64_t exit;
  uint64_t xr;
  uint32_t cnt;
  int16_t lbeta;
  uint16_t err;
};
static_assert(sizeof(SpanSum) == 24, "SpanSum layout");
struct SpanPick {
  uint64_t ent;
  uint64_t v0;
  uint32_t g0;
  int32_t beta0;
};
static_assert(sizeof(SpanPick) == 24, "SpanPick layout");
struct ChunkWalk {
  uint64_t cur;
  uint64_t v;
  uint32_t frontier;  // the span count: done
  uint32_t g;
  int32_t beta;       // INT32_MAX: unset
  uint32_t pad;
};
static_assert(sizeof(ChunkWalk) == 32, "ChunkWalk layout");
// Arguments of the index kernels (column cols[0]): span_first (per chunk
// its first span in the span arrays), groups (chunk, first span) of
// kSpansPerLane spans, per span cand / res / pick, per chunk walk and ev
// (its first event, record << 8 | code, ~0 none), the load's error flags.
struct XIndexArgs {
  const ColView *cols;
  const uint32_t *span_first;
  const uint2 *groups;
  uint64_t *cand;
  SpanSum *res;
  SpanPick *pick;
  ChunkWalk *walk;
  uint64_t *ev;
  uint32_t *err;
  uint32_t n_groups, nc;
};

// knn1 candidates of one chunk (Decoder::knn1, api/decoder.rs:469-533):
// `straddle` when the previous chunk ended inside a vector (its partial
// covered target[0 .. m1) from record prev_off of that chunk), then `k`
// whole vectors from record 0.  n_eff: the chunk's decoded records.
struct alignas(16) KnnChunk {
  uint64_t rec_off, prev_rec_off;
  uint32_t n_eff, prev_n_eff;
  uint32_t prev_off, m1;
  uint32_t straddle, k;
};
static_assert(sizeof(KnnChunk) == 40 || sizeof(KnnChunk) == 48, "KnnChunk layout");

// Decoder::matmul batches of one chunk (api/decoder.rs:606-656): they start
// at record off0 of the chunk, one data matrix apart.
struct alignas(16) DotChunk {
  uint64_t rec_off;
  uint32_t n_eff, off0;
};

// Postfix program of an Expr (decoder/expr.rs): kind 0 column `index`,
// 1 literal, 2 add, 3 sub, 4 mul, 5 div.
constexpr uint32_t kExprNodes = 32;
constexpr int kExprStack = 16;
struct ExprNode {
  int32_t kind;
  uint32_t index;
  double literal;
};
struct ExprProgram {
  ExprNode nodes[kExprNodes];
  uint32_t n;
};

// GPU encoder (BUFF, DeltaSprintz): per-chunk statistics (BUFF: a = min and
// b = max fixed-point value; DeltaSprintz: b = OR of the zigzagged deltas)
// and the layout the host derives from them (body offset in the file body,
// BUFF base / DeltaSprintz unused, bit width B or W, BUFF fractional bits).
struct EncStat {
  int64_t a, b;
};
struct EncChunk {
  uint64_t off;
  int64_t base;
  uint32_t bits, frac;
};

// One query of a group as the group kernels see it (query plans, host/
// plan.cpp): its arguments, its column, its scratch and what to run.
enum PlanFlags : uint32_t {
  kPfTiles = 1u << 0,    // record-tile pass (in-situ BUFF / DeltaSprintz, decoded records)
  kPfSegs = 1u << 1,     // fused decode + query over the column's segment jobs
  kPfPred = 1u << 2,     // evaluate the predicate (filter)
  kPfBits = 1u << 3,     // write the selected records' bits to a.out_bits
  kPfAggSel = 1u << 6,   // the aggregate runs over the predicate's selection (FILTER_SUM)
  kPfFold = 1u << 7,     // per-chunk results folded on the device into result->scalar
  kPfCompact = 1u << 8,  // materialize: compaction of a.values by a.out_bits into `out`
  kPfMask = 1u << 9,     // a bitmask was given
  kPfF64Sum = 1u << 10,  // the sum is over f64 records (double-double tile partials)
  kPfFieldTiles = 1u << 11,  // RLE queried per run: the tile pass walks field tiles (partials per field tile)
  kPfCpred = 1u << 12,       // the selection pass writes each chunk's ChunkPred (BUFF / DeltaSprintz)
};
constexpr uint32_t kPfAggShift = 4;  // bits 4-5: 0 none, 1 sum, 2 min, 3 max

struct alignas(16) PlanResult {
  double scalar;   // SUM / MIN / MAX / FILTER_SUM
  uint64_t n_out;  // MATERIALIZE / FILTER_MATERIALIZE through compaction
};

// The record-tile kernels of a group: BUFF codes, DeltaSprintz codes, the
// bitmask / decoded records (Uncompressed: the chunk body itself), and RLE
// runs.
enum SpanKind : int { kSpanBuff = 0, kSpanSprintz = 1, kSpanValues = 2, kSpanRle = 3 };
constexpr int kSpanKinds = 4;

struct alignas(16) PlanDesc {
  QueryArgs a;
  const ColView *view;
  const Job *span_jobs;  // the column's record-tile span jobs (Job::kind = tiles)
  double *chunk_sum;     // per chunk: (hi, lo) of an f64 sum, or (value, 0) to fold
  double *out;           // compaction output
  uint64_t *block;       // compaction: per-block counts, then offsets
  PlanResult *result;
  uint64_t words;        // ceil(N / 64)
  int64_t chunk_id;
  uint32_t n_chunks, n_span_jobs, n_seg_jobs, n_blocks;
  uint32_t flags;
  int32_t filter_mode;   // 0 select by the bitmask, 1 code domain (BUFF / DeltaSprintz) or f64, 2 f64
  uint32_t scheme;
  uint32_t pad;
};

// A fill of a group's scratch (k_group_fill): n words at p, word i = v[i % period].
struct alignas(16) GroupFill {
  uint64_t *p;
  uint64_t n;
  uint64_t v[4];
  uint32_t period;
  uint32_t pad0;
  uint64_t pad1;
};

// Device error flags (OR-ed by kernels).
enum DevError : uint32_t {
  kErrNone = 0,
  kErrOverrun = 1,      // a reader went past the chunk's bytes
  kErrCount = 2,        // terminator or run total disagrees with the record count
  kErrCorrupt = 4,      // field combination the reference panics on
  kNanEnd = 8,          // not an error: a NaN end marker before the footer's record
                        // count ends the chunk there (returned by the serial decoders
                        // with the marker's record index in XorState::end)
};

}  // namespace ebi
