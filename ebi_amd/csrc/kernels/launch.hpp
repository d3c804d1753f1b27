// Host-side launchers of the libebi_hip kernels (all asynchronous on `s`).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../device.hpp"

namespace ebi {
namespace launch {

hipError_t set_pow10(const double *p10, const double *n10);
void zero_u32(hipStream_t s, uint32_t *p, uint32_t n);

// Decode: tile jobs (plain, BUFF, Sprintz, RLE) in one persistent-style launch.
void tiles(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, uint32_t *ctr,
           uint32_t *err, uint32_t max_blocks);
// Lane-per-chunk serial XOR decode (Gorilla, Chimp, Elf); `index` also writes checkpoints.
void serial(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, uint32_t *err,
            bool write, bool index);
void chimp128(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, uint32_t *err,
              bool write);
// Checkpoint-segment XOR decode.
void segments(hipStream_t s, const ColView *cols, const Job *jobs, uint32_t n_jobs, uint32_t *err);
// Load-time tile index for DeltaSprintz and RLE chunks of column slot `col`.
void tile_index(hipStream_t s, const ColView *cols, uint32_t col, const Job *jobs, uint32_t n_jobs,
                uint32_t n_chunks, uint32_t *err);

}  // namespace launch
}  // namespace ebi
