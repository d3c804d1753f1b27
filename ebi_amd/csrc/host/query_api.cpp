// C ABI query entry points of libebi_hip: Decoder::{materialize, filter,
// filter_materialize, sum, min, max} (lemolatoon/ebi ebi/src/api/
// decoder.rs:202-439) over a resident column.
//
// Each call runs one query plan (plan.hpp) synchronously on the ctx stream:
// upload the optional bitmask, select chunks on the device (chunk_id and
// bitmask pruning, api/decoder.rs:216-226), run the in-situ kernels (BUFF,
// DeltaSprintz) or decode into HBM and run the record-tile kernels (other
// codecs), then copy the result back.  Per-chunk partial results are
// combined on the host in chunk order, as the Decoder loop does.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "../../../include/ebi_hip.h"
#include "column.hpp"
#include "internal.hpp"
#include "plan.hpp"

namespace ebi {

// Scratch of ebi_hip_query_noalloc: one array per GPU for the life of the
// process.  Host code never uses its name as an address (it names a
// host-side stand-in): noalloc_init looks up this GPU's copy once, with the
// device current, and the calls pass that address to the kernels.
constexpr uint64_t kNoAllocDeviceBytes = 64ull << 20;
constexpr uint64_t kNoAllocHostBytes = 4ull << 20;
constexpr int kNoAllocMaxDevices = 16;
static __device__ __align__(256) uint8_t scratch[kNoAllocDeviceBytes];
// host staging of the same calls (per-chunk results, chunk predicates, the
// launch's views and flags), one region per device
alignas(256) static uint8_t g_noalloc_host[kNoAllocMaxDevices][kNoAllocHostBytes];
static uint8_t *g_noalloc_dev[kNoAllocMaxDevices];
static std::mutex g_noalloc_mu;

int noalloc_init(ebi_hip_ctx *ctx) {
  if (ctx->device < 0 || ctx->device >= kNoAllocMaxDevices) return EBI_OK;  // the entry point refuses
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_dev[ctx->device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // ctx->device is current
    if (e != hipSuccess || !ptr) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
    g_noalloc_dev[ctx->device] = static_cast<uint8_t *>(ptr);
  }
  ctx->noalloc_dev = g_noalloc_dev[ctx->device];
  return EBI_OK;
}

namespace {

int ensure_pinned(ebi_hip_ctx *ctx, size_t bytes) {
  if (bytes <= ctx->plan_pin_cap) return 0;
  if (ctx->plan_pin) hipHostFree(ctx->plan_pin);
  ctx->plan_pin = nullptr;
  ctx->plan_pin_cap = 0;
  size_t want = 4096;
  while (want < bytes) want *= 2;
  if (hipHostMalloc((void **)&ctx->plan_pin, want, hipHostMallocDefault) != hipSuccess) return 1;
  ctx->plan_pin_cap = want;
  return 0;
}

int run_query(ebi_hip_ctx *ctx, const ebi_query &q) {
  Plan p;
  int rc = plan_init(p, q);
  if (rc) return fail(ctx, rc, p.error.empty() ? "invalid query arguments" : p.error);
  PlanGroup g;
  g.plans.push_back(&p);
  const PlanNeed need = plan_need(p);
  const uint64_t gb = group_block_bytes(1);
  if (ctx->plan_dev.ensure(need.dev + gb)) return fail(ctx, EBI_OOM, "query scratch allocation");
  if (ensure_pinned(ctx, need.pin + gb)) return fail(ctx, EBI_OOM, "pinned staging allocation");
  uint8_t *dev = (uint8_t *)ctx->plan_dev.p, *pin = ctx->plan_pin;
  group_bind(g, dev, pin, 1);
  plan_bind(p, dev + gb, pin + gb);
  plan_stage(p);
  hipEventRecord(ctx->ev0, ctx->stream);
  hipError_t e = group_enqueue(g, ctx->stream, ctx->cu_count, ctx->blocks_per_cu);
  if (e != hipSuccess) return hip_fail(ctx, e, "query launch");
  hipEventRecord(ctx->ev1, ctx->stream);
  if ((e = stream_wait(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "query");
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->last_ms = ms;
  if (g.h_flags[0]) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "query: device reported chunk error flags 0x%x", g.h_flags[0]);
    return fail(ctx, map_dev_err(g.h_flags[0]), buf);
  }
  if (plan_enqueue_values(p, ctx->stream, &e)) {
    if (e == hipSuccess) e = stream_wait(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "query result copy");
  }
  return plan_finish(p);
}

// ebi_hip_query without allocations: the plan's device scratch is this
// GPU's `scratch` array, its host staging the device's static region, and
// the bitmask / filter result move between the caller's buffers and the
// device directly.
int run_query_noalloc(ebi_hip_ctx *ctx, const ebi_query &q) {
  if (!ctx->noalloc_dev) return fail(ctx, EBI_UNSUPPORTED, "query_noalloc: no scratch on this device");
  Plan p;
  int rc = plan_init(p, q);
  p.direct = true;
  if (rc) return fail(ctx, rc, p.error.empty() ? "invalid query arguments" : p.error);
  const PlanNeed need = plan_need(p);
  const uint64_t gb = group_block_bytes(1);
  if (need.dev + gb > kNoAllocDeviceBytes || need.pin + gb > kNoAllocHostBytes) {
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "query_noalloc: the query needs %llu device and %llu host scratch bytes, above the fixed "
                  "%llu / %llu",
                  (unsigned long long)(need.dev + gb), (unsigned long long)(need.pin + gb),
                  (unsigned long long)kNoAllocDeviceBytes, (unsigned long long)kNoAllocHostBytes);
    return fail(ctx, EBI_INVALID_ARGUMENT, buf);
  }
  PlanGroup g;
  g.plans.push_back(&p);
  uint8_t *dev = ctx->noalloc_dev, *pin = g_noalloc_host[ctx->device];
  group_bind(g, dev, pin, 1);
  plan_bind(p, dev + gb, pin + gb);
  plan_stage(p);
  hipEventRecord(ctx->ev0, ctx->stream);
  hipError_t e = group_enqueue(g, ctx->stream, ctx->cu_count, ctx->blocks_per_cu);
  if (e != hipSuccess) return hip_fail(ctx, e, "query_noalloc launch");
  hipEventRecord(ctx->ev1, ctx->stream);
  if ((e = stream_wait(ctx->stream)) != hipSuccess) return hip_fail(ctx, e, "query_noalloc");
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->last_ms = ms;
  if (g.h_flags[0]) {
    char buf[96];
    std::snprintf(buf, sizeof buf, "query_noalloc: device reported chunk error flags 0x%x", g.h_flags[0]);
    return fail(ctx, map_dev_err(g.h_flags[0]), buf);
  }
  if (plan_enqueue_values(p, ctx->stream, &e)) {
    if (e == hipSuccess) e = stream_wait(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "query_noalloc result copy");
  }
  return plan_finish(p);
}

ebi_query make_query(int op, const ebi_hip_col *col, const ebi_pred *pred, const uint64_t *bitset,
                     int64_t chunk_id) {
  ebi_query q{};
  q.op = op;
  q.col = col;
  if (pred) q.pred = *pred;
  q.bitset = bitset;
  q.chunk_id = chunk_id;
  return q;
}

}  // namespace
}  // namespace ebi

using namespace ebi;

extern "C" int ebi_hip_query(ebi_hip_ctx *ctx, const ebi_query *q) {
  EBI_GUARD(ctx);
  if (!q) return EBI_INVALID_ARGUMENT;
  return run_query(ctx, *q);
}

extern "C" int ebi_hip_query_noalloc(ebi_hip_ctx *ctx, const ebi_query *q) {
  EBI_GUARD(ctx);
  if (!q) return EBI_INVALID_ARGUMENT;
  return run_query_noalloc(ctx, *q);
}

extern "C" int ebi_hip_sum(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset, int64_t chunk_id,
                           double *out) {
  EBI_GUARD(ctx);
  ebi_query q = make_query(EBI_OP_SUM, col, nullptr, bitset, chunk_id);
  q.out_scalar = out;
  return run_query(ctx, q);
}
extern "C" int ebi_hip_min(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset, int64_t chunk_id,
                           double *out) {
  EBI_GUARD(ctx);
  ebi_query q = make_query(EBI_OP_MIN, col, nullptr, bitset, chunk_id);
  q.out_scalar = out;
  return run_query(ctx, q);
}
extern "C" int ebi_hip_max(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset, int64_t chunk_id,
                           double *out) {
  EBI_GUARD(ctx);
  ebi_query q = make_query(EBI_OP_MAX, col, nullptr, bitset, chunk_id);
  q.out_scalar = out;
  return run_query(ctx, q);
}

extern "C" int ebi_hip_filter(ebi_hip_ctx *ctx, const ebi_hip_col *col, const ebi_pred *pred, const uint64_t *bitset,
                              int64_t chunk_id, uint64_t *out_bitset) {
  EBI_GUARD(ctx);
  if (!pred) return EBI_INVALID_ARGUMENT;
  ebi_query q = make_query(EBI_OP_FILTER, col, pred, bitset, chunk_id);
  q.out_bitset = out_bitset;
  return run_query(ctx, q);
}

extern "C" int ebi_hip_materialize(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset,
                                   int64_t chunk_id, double *out, size_t *n_out) {
  EBI_GUARD(ctx);
  if (n_out) *n_out = 0;
  ebi_query q = make_query(EBI_OP_MATERIALIZE, col, nullptr, bitset, chunk_id);
  q.out_values = out;
  q.n_out = n_out;
  return run_query(ctx, q);
}

extern "C" int ebi_hip_filter_materialize(ebi_hip_ctx *ctx, const ebi_hip_col *col, const ebi_pred *pred,
                                          const uint64_t *bitset, int64_t chunk_id, double *out, size_t *n_out) {
  EBI_GUARD(ctx);
  if (!pred) return EBI_INVALID_ARGUMENT;
  if (n_out) *n_out = 0;
  ebi_query q = make_query(EBI_OP_FILTER_MATERIALIZE, col, pred, bitset, chunk_id);
  q.out_values = out;
  q.n_out = n_out;
  return run_query(ctx, q);
}
