// Host-side launchers of the libebi_hip kernels (all asynchronous on `s`).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../device.hpp"

namespace ebi {
namespace launch {

hipError_t set_pow10(const double *p10, const double *n10);
void zero_u32(hipStream_t s, uint32_t *p, uint32_t n);

// Persistent decode of n_cols columns: k_decode over their tile jobs, then
// k_decode_seg over their checkpoint-segment jobs (Chimp128: parse and
// values in one pass) (tile_prefix / seg_prefix:
// n_cols + 1 entries each, the job prefix of each list; either may be null
// when its job count is 0).  Grids of at most max_blocks blocks; `mid`
// (nullable) is recorded between the two kernels.  mixed: when both lists
// are non-empty, one k_decode_mix launch over both instead (`mid` recorded
// after it).
// codecs: bit c set when some column has codec c (one k_decode launch per
// tile codec present; 0: every tile codec).
// side (nullable, with the events fork and join): when there are segment jobs
// and Uncompressed columns, their copy kernel runs on `side` with side_blocks
// blocks, forked after the other tile kernels and joined back into s after
// k_decode_seg, so it fills the wave slots and VGPRs the segment kernel leaves
// free (it needs no LDS); side_blocks stays small enough that the segment
// kernel's 4 blocks per CU still fit beside it.
constexpr uint32_t kTileCodecs = (1u << kCodecPlain) | (1u << kCodecRLE) | (1u << kCodecBUFF) | (1u << kCodecSprintz);
void decode(hipStream_t s, const ColView *cols, uint32_t n_cols, const uint32_t *tile_prefix, uint32_t n_tile_jobs,
            const uint32_t *seg_prefix, uint32_t n_seg_jobs, uint32_t *err, uint32_t max_blocks,
            hipEvent_t mid = nullptr, bool mixed = false, uint32_t codecs = 0, hipStream_t side = nullptr,
            hipEvent_t fork = nullptr, hipEvent_t join = nullptr, uint32_t side_blocks = 0);
// Lane-per-chunk serial XOR decode (Gorilla, Chimp, Elf); `index` writes
// checkpoints.  prefix is over n_serial_jobs; `single` (device, nullable)
// runs one explicit job on cols[0] instead.
void serial(hipStream_t s, const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t n_jobs,
            const Job *single, uint32_t *err, bool write, bool index);
// Load-time ring checkpoints of the Chimp128 column cols[0] (decoded into
// cols[0].out): per segment job the values of the 128 records before it.
void c128_ring(hipStream_t s, const ColView *cols, uint32_t n_jobs, uint64_t *ring);
// One step of the wavefront ring build: job `step` of every chunk of the
// Chimp128 column cols[0] (job_first: nc + 1 prefix of its segment jobs).
void c128_ring_step(hipStream_t s, const ColView *cols, const uint32_t *job_first, uint32_t nc, uint32_t step,
                    uint64_t *ring, uint32_t *err);
void chimp128(hipStream_t s, const ColView *cols, const uint32_t *prefix, uint32_t n_cols, uint32_t n_jobs,
              const Job *single, uint32_t *err, bool write);
// Load-time lane-interleaved segment streams of the XOR column cols[0]:
// ilv_rows writes the row count of each segment job; ilv_copy fills the
// rows (each job's first row in its Job::col).
void ilv_rows(hipStream_t s, const ColView *cols, uint32_t n_jobs, uint32_t *rows);
void ilv_copy(hipStream_t s, const ColView *cols, uint32_t n_jobs, const uint32_t *rows, uint4 *ilv);
// Load-time parallel index of the XOR column cols[0] (kernels/xindex.hpp):
// initial state, candidates (Gorilla: from its walkers' frontiers on), the
// per-chunk walkers up to span `stop` (blocks: Gorilla's walker blocks),
// then checkpoints and per-chunk events (NaN ends into chunk_end, errors
// into chunk_err and a.err).
void xindex_init(hipStream_t s, const XIndexArgs &a, uint32_t n_spans);
void xindex_cands(hipStream_t s, const XIndexArgs &a);
void xindex_walk(hipStream_t s, const XIndexArgs &a, uint32_t stop, bool blocks);
void xindex_emit(hipStream_t s, const XIndexArgs &a);
// Load-time tile index for DeltaSprintz and RLE chunks of column slot `col`.
void tile_index(hipStream_t s, const ColView *cols, uint32_t col, const Job *jobs, uint32_t n_jobs,
                uint32_t n_chunks, uint32_t *err);

// Queries (record-tile jobs; single-column views at cols[0]).
void chunk_select(hipStream_t s, const ColView *cols, const uint64_t *mask, int64_t chunk_id, uint32_t *sel,
                  uint32_t n_chunks);
void filter(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, const QueryArgs &a, int mode);
void aggregate(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, const QueryArgs &a,
               uint32_t n_chunks);
// The kernels of a group of query plans (PlanDesc array in device memory,
// flat work items mapped to plans by prefix arrays of n_plans + 1 entries):
// scratch fills and chunk selection in one launch (items: chunks), the
// record-tile span pass of one SpanKind (items: span jobs), the fused decode + query over segment jobs (items:
// segment jobs; persistent grid of at most max_blocks; one launch per codec
// in `codecs`, bit c for codec c, 0: every segment codec), the per-chunk fold
// and the final fold (items: chunks), and stream compaction (items: 32,768-
// record blocks).
void group_prep(hipStream_t s, const GroupFill *fills, uint32_t n_fills, uint64_t max_words, const PlanDesc *plans,
                const uint32_t *prefix, uint32_t n_plans, uint32_t n_items);
void group_tiles(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_jobs,
                 int kind);
constexpr uint32_t kSegCodecs = (1u << kCodecGorilla) | (1u << kCodecChimp) | (1u << kCodecChimp128) |
                                (1u << kCodecElfOnChimp) | (1u << kCodecElf);
void query_seg(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_jobs,
               uint32_t *err, uint32_t max_blocks, uint32_t codecs = 0);
void group_fold(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_items,
                bool any);
void group_compact(hipStream_t s, const PlanDesc *plans, const uint32_t *prefix, uint32_t n_plans, uint32_t n_blocks);
void chunk_dd(hipStream_t s, const ColView *cols, const uint32_t *sel, const double *partials, double *chunk_sum,
              uint32_t n_chunks);
uint32_t compact_blocks(uint64_t n_words);
void compact(hipStream_t s, const uint64_t *bits, uint64_t n_words, const double *values, double *out,
             uint64_t *block_scratch, uint64_t *n_out);

// Analytics over decoded columns (analytics.hip).  knn: candidate distances
// to d[], then red[0] = min distance key, red[1] = first candidate at it,
// red[2] = 1 + last `<=` candidate at it (0: none).
void knn(hipStream_t s, const double *values, const KnnChunk *chunks, uint32_t nc, const uint64_t *cand0,
         uint64_t n_cand, const double *target, uint32_t L, double *d, unsigned long long *red);
void dot(hipStream_t s, const double *values, const DotChunk *chunks, uint32_t nc, const uint64_t *batch0,
         uint64_t n_batch, uint32_t stride, uint32_t rows, uint32_t data_cols, const double *target, uint32_t k,
         uint32_t cols, double *out);
void dist_one(hipStream_t s, const double *values, uint64_t rec_off, uint32_t n_eff, uint32_t off,
              const double *target, uint32_t len, int dot, double *out);
void chunk_count(hipStream_t s, const DevChunk *chunks, uint32_t nc, const uint64_t *bits, uint64_t *counts);
void expr(hipStream_t s, const ExprProgram &prog, const double *const *cols, const uint64_t *col_off, uint32_t nc,
          const uint64_t *len, const uint64_t *out_off, const uint2 *tiles, uint32_t n_tiles, int kind,
          double *partials, unsigned long long *acc, double *out);

// GPU encoders (encode.hip): per-chunk statistics of n values (chunks of
// per_chunk records), then the chunk bodies packed into `file` at each
// EncChunk::off.  codec kCodecBUFF or kCodecSprintz.
void enc_stats(hipStream_t s, const double *v, uint64_t n, uint64_t per_chunk, uint64_t nc, int codec, int32_t frac,
               double scale, EncStat *stats);
void enc_pack(hipStream_t s, const double *v, uint64_t n, uint64_t per_chunk, uint64_t nc, int codec,
              const EncChunk *chunks, double scale, uint64_t scale_bits, uint8_t *file);

// FP64 DGEMM on the matrix cores: C_b = A_b * B_b, row-major, b < batch,
// element strides sA / sB / sC between batches (0 = shared operand).
hipError_t dgemm(hipStream_t s, const double *A, const double *B, double *C, uint32_t m, uint32_t n, uint32_t k,
                 uint64_t batch, uint64_t sA, uint64_t sB, uint64_t sC);

}  // namespace launch
}  // namespace ebi
