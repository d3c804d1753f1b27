// Helpers shared by the host translation units of libebi_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "column.hpp"

// Entry-point prologue: null check, serialise on the ctx, clear the error.
#define EBI_GUARD(ctx)                                   \
  if (!(ctx)) return EBI_INVALID_ARGUMENT;               \
  std::lock_guard<std::recursive_mutex> _lk((ctx)->mu); \
  (ctx)->last_error.clear();

namespace ebi {

int fail(ebi_hip_ctx *ctx, int code, const std::string &msg);
int hip_fail(ebi_hip_ctx *ctx, hipError_t e, const char *what);
int map_dev_err(uint32_t flags);
bool is_xor_codec(int scheme);
// Waits for the work queued on `s` by polling it (the synchronous entry
// points are latency-bound: a blocking wait adds the host thread's wake-up
// time to every call); returns the first error the stream reports.
hipError_t stream_wait(hipStream_t s);
// The ColView of `c` with output base `out`.
ColView view_of(const Column &c, double *out);
// Decode tile jobs of a tile-codec column with up to per_job tiles (RLE:
// output steps of one field tile) per job; first (nullable) gets each
// chunk's first job, nc + 1 entries.  Chunks rejected at load get none
// (RLE) or are skipped by the kernel.
std::vector<Job> tile_jobs_of(const Column &c, uint32_t per_job, std::vector<uint32_t> *first);
// Uploads the single-column ColView of `c` with output base `out` and waits.
int upload_view(ebi_hip_ctx *ctx, Column &c, double *out);
// Copies the device error word back (synchronising the stream) and maps it.
int sync_and_check(ebi_hip_ctx *ctx, const char *what);
// Fails if a selected chunk was rejected by the load-time index pass.
int check_chunk_errors(ebi_hip_ctx *ctx, const Column &c, int64_t chunk_id);
// Queues the kernels that decode every chunk of `c` into the view's output.
int enqueue_decode(ebi_hip_ctx *ctx, Column &c);
// Looks up the ctx device's scratch of ebi_hip_query_noalloc (device current).
int noalloc_init(ebi_hip_ctx *ctx);

}  // namespace ebi
