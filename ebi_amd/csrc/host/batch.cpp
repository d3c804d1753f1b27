// Prepared batched decode: a fixed set of resident columns decoded into fixed
// device outputs, replayed many times.  The job-space arguments are uploaded
// once; an autotuner times candidate persistent grid sizes on this very
// batch and keeps the fastest; the launch sequence is captured in a hipGraph
// and replayed with one launch.  No counterpart in lemolatoon/ebi,
// whose Decoder decodes chunk after chunk on the CPU (api/decoder.rs:202-241).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../../include/ebi_hip.h"
#include "../kernels/launch.hpp"
#include "column.hpp"
#include "internal.hpp"

using namespace ebi;

struct ebi_hip_batch {
  std::vector<Column *> cols;
  std::vector<double *> outs;
  DevBuf args;     // prefix[n + 1] | prefix_seg[n + 1] | pad | ColView[n]
  uint64_t pre_bytes = 0;
  uint32_t n_cols = 0, jobs = 0, segs = 0;
  ebi_batch_config cfg{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

const ColView *views_of(const ebi_hip_batch *b) {
  return reinterpret_cast<const ColView *>(static_cast<uint8_t *>(b->args.p) + b->pre_bytes);
}
const uint32_t *prefix_of(const ebi_hip_batch *b) { return static_cast<const uint32_t *>(b->args.p); }

// Issues the batch on ctx->stream with configuration c.  timed: record the
// phase events (eager runs only; a captured sequence carries no timing).
void enqueue(ebi_hip_ctx *ctx, ebi_hip_batch *b, const ebi_batch_config &c, bool timed) {
  hipStream_t s = ctx->stream;
  launch::zero_u32(s, ctx->d_flags, 2);
  if (timed) hipEventRecord(ctx->ev0, s);
  launch::decode(s, views_of(b), b->n_cols, prefix_of(b), b->jobs, prefix_of(b) + (b->n_cols + 1), b->segs,
                 ctx->d_flags, (uint32_t)ctx->cu_count * c.blocks_per_cu, timed ? ctx->ev_seg : nullptr,
                 c.mixed != 0);
  if (timed) hipEventRecord(ctx->ev_mid, s);
  if (timed) hipEventRecord(ctx->ev1, s);
}

// phase 0: k_decode (or k_decode_mix), 1: none (kept 0), 2: k_decode_seg (0 when mixed)
void read_times(ebi_hip_ctx *ctx) {
  float a = 0, c = 0, g = 0, t = 0;
  hipEventElapsedTime(&a, ctx->ev0, ctx->ev_seg);
  hipEventElapsedTime(&g, ctx->ev_seg, ctx->ev_mid);
  hipEventElapsedTime(&c, ctx->ev_mid, ctx->ev1);
  hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  ctx->phase_ms[0] = a;
  ctx->phase_ms[1] = c;
  ctx->phase_ms[2] = g;
  ctx->phase_ms[3] = 0.0;
  ctx->last_ms = t;
}

// Median device time of `reps` eager runs of configuration c (after one
// warm-up); < 0 on error.
double time_config(ebi_hip_ctx *ctx, ebi_hip_batch *b, const ebi_batch_config &c, int reps) {
  std::vector<double> ms;
  for (int r = 0; r <= reps; ++r) {
    enqueue(ctx, b, c, true);
    if (sync_and_check(ctx, "batch tuning")) return -1.0;
    read_times(ctx);
    if (r) ms.push_back(ctx->last_ms);
  }
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

int capture(ebi_hip_ctx *ctx, ebi_hip_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch capture begin");
  enqueue(ctx, b, b->cfg, false);
  e = hipStreamEndCapture(ctx->stream, &b->graph);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch capture end");
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch graph instantiate");
  return EBI_OK;
}

void destroy(ebi_hip_batch *b) {
  for (Column *c : b->cols) c->users.fetch_sub(1);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->args.release();
  delete b;
}

}  // namespace

extern "C" int ebi_hip_batch_create(ebi_hip_ctx *ctx, const ebi_hip_col *const *cols, size_t n_cols,
                                    double *const *d_outs, int autotune, ebi_hip_batch **out) {
  EBI_GUARD(ctx);
  if (!cols || !d_outs || !out || n_cols == 0 || n_cols > (1u << 16)) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  int rc;
  for (size_t i = 0; i < n_cols; ++i) {
    if (!cols[i] || (!d_outs[i] && cols[i]->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
    if ((rc = check_chunk_errors(ctx, cols[i]->c, -1))) return rc;
  }
  ebi_hip_batch *b = new (std::nothrow) ebi_hip_batch();
  if (!b) return EBI_OOM;
  b->n_cols = (uint32_t)n_cols;
  b->pre_bytes = ((2 * (n_cols + 1) * 4 + 255) / 256) * 256;
  std::vector<uint8_t> host(b->pre_bytes + n_cols * sizeof(ColView));
  uint32_t *pre = reinterpret_cast<uint32_t *>(host.data()), *preseg = pre + (n_cols + 1);
  ColView *views = reinterpret_cast<ColView *>(host.data() + b->pre_bytes);
  uint64_t jobs = 0, segs = 0;
  for (size_t i = 0; i < n_cols; ++i) {
    Column &c = const_cast<ebi_hip_col *>(cols[i])->c;
    c.users.fetch_add(1);  // the captured graph holds the column's addresses
    b->cols.push_back(&c);
    b->outs.push_back(d_outs[i]);
    pre[i] = (uint32_t)jobs;
    preseg[i] = (uint32_t)segs;
    views[i] = view_of(c, d_outs[i]);
    jobs += c.n_tile_jobs;
    segs += c.n_seg_jobs;
  }
  pre[n_cols] = (uint32_t)jobs;
  preseg[n_cols] = (uint32_t)segs;
  if (jobs >= (1ull << 32) || segs >= (1ull << 32)) {
    destroy(b);
    return fail(ctx, EBI_UNSUPPORTED, "batch too large");
  }
  b->jobs = (uint32_t)jobs;
  b->segs = (uint32_t)segs;
  hipError_t e;
  if (b->args.ensure(host.size()) ||
      (e = hipMemcpy(b->args.p, host.data(), host.size(), hipMemcpyHostToDevice)) != hipSuccess) {
    destroy(b);
    return fail(ctx, EBI_OOM, "batch arguments");
  }
  b->cfg.blocks_per_cu = ctx->blocks_per_cu;
  b->cfg.mixed = ctx->batch_mixed && b->jobs && b->segs ? 1u : 0u;
  if (autotune) {
    // persistent grid size of the decode kernels, and (with both job kinds)
    // one mixed launch or the two kernels in turn
    double best = time_config(ctx, b, b->cfg, 2);
    if (best < 0) {
      destroy(b);
      return EBI_HIP;
    }
    auto consider = [&](ebi_batch_config c) {
      const double t = time_config(ctx, b, c, 2);
      if (t >= 0 && t < best) {
        best = t;
        b->cfg = c;
      }
    };
    const ebi_batch_config first = b->cfg;
    const uint32_t n_mixed = b->jobs && b->segs ? 2u : 1u;
    for (uint32_t mixed = 0; mixed < n_mixed; ++mixed)
      for (uint32_t bpc : {4u, 8u, 16u}) {
        ebi_batch_config c = first;
        c.blocks_per_cu = bpc;
        c.mixed = mixed;
        if (bpc != first.blocks_per_cu || mixed != first.mixed) consider(c);
      }
    b->cfg.tuned_ms = best;
  }
  if ((rc = capture(ctx, b))) {
    destroy(b);
    return rc;
  }
  *out = b;
  return EBI_OK;
}

extern "C" int ebi_hip_batch_run(ebi_hip_ctx *ctx, ebi_hip_batch *b, int eager) {
  EBI_GUARD(ctx);
  if (!b) return EBI_INVALID_ARGUMENT;
  if (eager) {
    enqueue(ctx, b, b->cfg, true);
    int rc = sync_and_check(ctx, "batch run");
    if (rc) return rc;
    read_times(ctx);
    return EBI_OK;
  }
  hipEventRecord(ctx->ev0, ctx->stream);
  hipError_t e = hipGraphLaunch(b->exec, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch graph launch");
  hipEventRecord(ctx->ev1, ctx->stream);
  int rc = sync_and_check(ctx, "batch run");
  if (rc) return rc;
  float t = 0;
  hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  ctx->last_ms = t;
  ctx->phase_ms[0] = ctx->phase_ms[1] = ctx->phase_ms[2] = ctx->phase_ms[3] = 0.0;
  return EBI_OK;
}

extern "C" int ebi_hip_batch_config_get(const ebi_hip_batch *b, ebi_batch_config *cfg) {
  if (!b || !cfg) return EBI_INVALID_ARGUMENT;
  *cfg = b->cfg;
  return EBI_OK;
}

extern "C" int ebi_hip_batch_destroy(ebi_hip_ctx *ctx, ebi_hip_batch *b) {
  EBI_GUARD(ctx);
  if (!b) return EBI_INVALID_ARGUMENT;
  hipStreamSynchronize(ctx->stream);
  destroy(b);
  return EBI_OK;
}
