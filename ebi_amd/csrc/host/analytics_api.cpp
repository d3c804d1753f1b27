// C ABI of the analytics over resident columns: the reference's ML queries
// (Decoder::knn1 and Decoder::matmul, api/decoder.rs:469-656; the chunk
// readers' distance_squared / dot_product, decoder/chunk_reader.rs:233-260
// over decoder/query.rs:157-192) and multi-column aggregations of
// arithmetic expressions (Aggregation::compute / compute_multiple,
// api/decoder.rs:835-917; Expr, decoder/expr.rs:67-123).
//
// Each call decodes the columns it reads into HBM (the batched decode
// kernels, the column's precision applied as the reference's chunk readers
// apply it), then runs the kernels of kernels/analytics.hip.  The reference's
// loop structure -- which records each vector, dot product or expression
// record reads -- is worked out on the host per chunk (chunk geometry from
// the footer, decoded lengths from the load-time passes) and handed to the
// kernels as small per-chunk tables.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../../include/ebi_hip.h"
#include "../kernels/launch.hpp"
#include "column.hpp"
#include "internal.hpp"

namespace ebi {
namespace {

// Device allocations of one call, freed when it returns (after the stream
// has drained).
struct Scratch {
  ebi_hip_ctx *ctx;
  std::vector<void *> ptrs;
  explicit Scratch(ebi_hip_ctx *c) : ctx(c) {}
  ~Scratch() {
    hipStreamSynchronize(ctx->stream);
    for (void *p : ptrs) hipFree(p);
  }
  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return static_cast<T *>(p);
  }
  template <class T>
  T *upload(const std::vector<T> &v) {
    T *d = alloc<T>(v.size());
    if (d && !v.empty() && hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
      return nullptr;
    return d;
  }
};

// Decodes `c` into `out` (N doubles) and waits; the column's precision applies.
int decode_column(ebi_hip_ctx *ctx, Column &c, double *out) {
  int rc = check_chunk_errors(ctx, c, -1);
  if (rc) return rc;
  if ((rc = upload_view(ctx, c, out))) return rc;
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  if ((rc = enqueue_decode(ctx, c))) return rc;
  return sync_and_check(ctx, "decode");
}

uint64_t footer_n(const Column &c, size_t k) { return c.chunks[k].geom.number_of_records; }

double key_to_f64(uint64_t key) {
  const uint64_t b = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
  double v;
  std::memcpy(&v, &b, 8);
  return v;
}

// Two-sum accumulation of double-double partials in a fixed order.
void dd_acc(double &hi, double &lo, double b) {
  const double s = hi + b;
  if (!std::isfinite(s)) {
    hi = s;
    lo = 0.0;
    return;
  }
  const double v = s - hi;
  lo += (hi - (s - v)) + (b - v);
  hi = s;
}

}  // namespace
}  // namespace ebi

using namespace ebi;

extern "C" int ebi_hip_knn1(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const double *target, size_t len,
                            ebi_knn_result *out) {
  EBI_GUARD(ctx);
  if (!colc || !out || (!target && len)) return EBI_INVALID_ARGUMENT;
  // an empty target never advances the reference's loop (api/decoder.rs:506-522)
  if (len == 0 || len > UINT32_MAX) return fail(ctx, EBI_INVALID_ARGUMENT, "knn1: target length must be 1..2^32-1");
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  const uint64_t N = c.meta.number_of_records;
  const uint32_t L = (uint32_t)len;
  const size_t nc = c.chunks.size();
  // per chunk: a completion of the vector the previous chunk ended inside
  // (its partial covered target[0 .. m1) from record k * L), then the whole
  // vectors from record 0 (the reference does not skip the completion's
  // records: api/decoder.rs:503-505 adds the reset remaining count, 0)
  std::vector<KnnChunk> kc(nc);
  std::vector<uint64_t> cand0(nc + 1, 0);
  uint64_t total = 0;
  for (size_t k = 0; k < nc; ++k) {
    KnnChunk &e = kc[k];
    e = KnnChunk{};
    e.rec_off = c.chunks[k].dev.rec_off;
    e.n_eff = c.chunks[k].dev.n;
    if (k > 0) {
      const uint64_t np = footer_n(c, k - 1);
      if (np % L) {
        e.straddle = 1;
        e.prev_rec_off = c.chunks[k - 1].dev.rec_off;
        e.prev_n_eff = c.chunks[k - 1].dev.n;
        e.prev_off = (uint32_t)(np / L * L);
        e.m1 = (uint32_t)(np % L);
      }
    }
    e.k = (uint32_t)(footer_n(c, k) / L);
    cand0[k] = total;
    total += e.straddle + e.k;
  }
  cand0[nc] = total;
  Scratch sc(ctx);
  double *values = sc.alloc<double>(N);
  double *d = sc.alloc<double>(total);
  unsigned long long *red = sc.alloc<unsigned long long>(4);
  if (!values || !d || !red) return fail(ctx, EBI_OOM, "knn1 scratch");
  int rc = decode_column(ctx, c, values);
  if (rc) return rc;
  std::vector<double> tv(target, target + len);
  const KnnChunk *dk = sc.upload(kc);
  const uint64_t *dc0 = sc.upload(cand0);
  const double *dt = sc.upload(tv);
  if (!dk || !dc0 || !dt) return fail(ctx, EBI_OOM, "knn1 tables");
  launch::knn(ctx->stream, values, dk, (uint32_t)nc, dc0, total, dt, L, d, red);
  unsigned long long h[3] = {0, 0, 0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, red, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "knn1");
  // The reference keeps the last candidate it accepted: a `<=` candidate at
  // or below the running minimum, a `<` (straddling) one strictly below it.
  // With D the smallest distance (NaN never accepted): the first candidate
  // at D is accepted when D < INFINITY, and every later `<=` candidate at D.
  out->index = UINT64_MAX;
  out->distance = INFINITY;
  if (h[0] == ~0ull) return EBI_OK;
  const double D = key_to_f64(h[0]);
  uint64_t idx = UINT64_MAX;
  if (D < INFINITY) idx = std::max<uint64_t>(h[1], h[2] ? h[2] - 1 : 0);
  else if (h[2]) idx = h[2] - 1;
  if (idx == UINT64_MAX) return EBI_OK;
  out->index = idx;
  out->distance = D;
  return EBI_OK;
}

static int chunk_fold(ebi_hip_ctx *ctx, const ebi_hip_col *colc, uint64_t chunk_id, size_t offset,
                      const double *target, size_t len, int dot, double *out) {
  if (!colc || !out || (!target && len) || len > UINT32_MAX || offset > UINT32_MAX) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  if (chunk_id >= c.chunks.size()) return fail(ctx, EBI_INVALID_ARGUMENT, "chunk id out of range");
  Scratch sc(ctx);
  double *values = sc.alloc<double>(c.meta.number_of_records);
  double *res = sc.alloc<double>(1);
  if (!values || !res) return fail(ctx, EBI_OOM, "scratch");
  int rc = decode_column(ctx, c, values);
  if (rc) return rc;
  std::vector<double> tv(target, target + len);
  const double *dt = sc.upload(tv);
  if (!dt) return fail(ctx, EBI_OOM, "target upload");
  const DevChunk &dc = c.chunks[chunk_id].dev;
  launch::dist_one(ctx->stream, values, dc.rec_off, dc.n, (uint32_t)offset, dt, (uint32_t)len, dot, res);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, res, 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, dot ? "dot_product" : "distance_squared");
  return EBI_OK;
}

extern "C" int ebi_hip_distance_squared(ebi_hip_ctx *ctx, const ebi_hip_col *col, uint64_t chunk_id,
                                        size_t offset_in_chunk, const double *target, size_t len, double *out) {
  EBI_GUARD(ctx);
  return chunk_fold(ctx, col, chunk_id, offset_in_chunk, target, len, 0, out);
}

extern "C" int ebi_hip_dot_product(ebi_hip_ctx *ctx, const ebi_hip_col *col, uint64_t chunk_id,
                                   size_t offset_in_chunk, const double *target, size_t len, double *out) {
  EBI_GUARD(ctx);
  return chunk_fold(ctx, col, chunk_id, offset_in_chunk, target, len, 1, out);
}

extern "C" int ebi_hip_matmul_dot(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const double *target, size_t target_rows,
                                  size_t target_cols, size_t data_rows, size_t data_cols, double *out) {
  EBI_GUARD(ctx);
  if (!colc || (!target && target_rows * target_cols) || !out) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  const uint64_t N = c.meta.number_of_records;
  // preconditions of Decoder::matmul (api/decoder.rs:613-620); a zero-sized
  // data matrix divides by zero there
  if (!data_rows || !data_cols) return fail(ctx, EBI_INVALID_ARGUMENT, "matmul: empty data matrix");
  if (data_cols != target_rows || N % (data_rows * data_cols) != 0)
    return fail(ctx, EBI_PRECONDITION, "matmul: matrix shapes do not match the column");
  if (data_rows > UINT32_MAX || data_cols > UINT32_MAX || target_cols > UINT32_MAX || target_rows > UINT32_MAX)
    return fail(ctx, EBI_UNSUPPORTED, "matmul: dimension above 2^32");
  const uint64_t S = (uint64_t)data_rows * data_cols;
  if (S > UINT32_MAX) return fail(ctx, EBI_UNSUPPORTED, "matmul: data matrix above 2^32 records");
  const uint64_t batches = N / S;
  // The reference's walk (api/decoder.rs:629-650): a batch starts where the
  // previous one ended, or at record 0 of the next chunk once that offset
  // reaches the chunk's record count (checked once per batch).  So chunk c
  // holds ceil(n_c / S) batches; an empty chunk after the first takes one
  // (an empty dot), the first none.
  const size_t nc = c.chunks.size();
  std::vector<DotChunk> dch;
  std::vector<uint64_t> b0;
  uint64_t done = 0;
  for (size_t k = 0; k < nc && done < batches; ++k) {
    const uint64_t n = footer_n(c, k);
    const uint64_t nb = n ? (n + S - 1) / S : (k ? 1 : 0);
    if (!nb) continue;
    dch.push_back(DotChunk{c.chunks[k].dev.rec_off, c.chunks[k].dev.n, 0});
    b0.push_back(done);
    done += std::min(nb, batches - done);
  }
  if (done < batches)  // the reference runs out of chunks (an out-of-range chunk id)
    return fail(ctx, EBI_PRECONDITION, "matmul: the data matrices overrun the column's chunks");
  b0.push_back(done);
  Scratch sc(ctx);
  double *values = sc.alloc<double>(N);
  const uint64_t n_out = batches * data_rows * target_cols;
  double *dout = sc.alloc<double>(n_out);
  if (!values || !dout) return fail(ctx, EBI_OOM, "matmul scratch");
  int rc = decode_column(ctx, c, values);
  if (rc) return rc;
  std::vector<double> tv(target, target + target_rows * target_cols);
  const double *dt = sc.upload(tv);
  const DotChunk *dd = sc.upload(dch);
  const uint64_t *db = sc.upload(b0);
  if (!dt || !dd || !db) return fail(ctx, EBI_OOM, "matmul tables");
  launch::dot(ctx->stream, values, dd, (uint32_t)dch.size(), db, batches, (uint32_t)S, (uint32_t)data_rows,
              (uint32_t)data_cols, dt, (uint32_t)target_rows, (uint32_t)target_cols, dout);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && n_out) e = hipMemcpyAsync(out, dout, n_out * 8, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "matmul");
  return EBI_OK;
}

namespace {

// Postfix program checks (decoder/expr.rs:67-99): a literal alone and a
// binary node of two literals are PreconditionsNotMet; column indexes must
// name a decoder.  *only_col: the program is one column (the in-situ path).
int check_program(const ebi_aggregation &a, size_t n_cols, ExprProgram *prog, int *only_col, std::vector<int> &used) {
  if (!a.nodes || a.n_nodes == 0 || a.n_nodes > kExprNodes) return EBI_INVALID_ARGUMENT;
  std::vector<int> lit;  // per stack entry: 1 if it is a bare literal
  prog->n = a.n_nodes;
  for (uint32_t i = 0; i < a.n_nodes; ++i) {
    const ebi_expr_node &e = a.nodes[i];
    ExprNode &p = prog->nodes[i];
    p.kind = e.kind;
    p.index = e.index;
    p.literal = e.literal;
    if (e.kind == EBI_EXPR_INDEX) {
      if (e.index >= n_cols) return EBI_INVALID_ARGUMENT;
      used[e.index] = 1;
      lit.push_back(0);
    } else if (e.kind == EBI_EXPR_LITERAL) {
      lit.push_back(1);
    } else if (e.kind >= EBI_EXPR_ADD && e.kind <= EBI_EXPR_DIV) {
      if (lit.size() < 2) return EBI_INVALID_ARGUMENT;
      const int b = lit.back();
      lit.pop_back();
      if (b && lit.back()) return EBI_PRECONDITION;
      lit.back() = 0;
    } else {
      return EBI_INVALID_ARGUMENT;
    }
    if (lit.size() > (size_t)kExprStack) return EBI_INVALID_ARGUMENT;
  }
  if (lit.size() != 1) return EBI_INVALID_ARGUMENT;
  if (lit[0]) return EBI_PRECONDITION;
  *only_col = a.n_nodes == 1 ? (int)a.nodes[0].index : -1;
  return EBI_OK;
}

}  // namespace

extern "C" int ebi_hip_aggregate(ebi_hip_ctx *ctx, const ebi_hip_col *const *cols, size_t n_cols,
                                 ebi_aggregation *aggs, size_t n_aggs, const uint64_t *bitset) {
  EBI_GUARD(ctx);
  if (!cols || !n_cols || !aggs || !n_aggs) return EBI_INVALID_ARGUMENT;
  for (size_t i = 0; i < n_cols; ++i)
    if (!cols[i]) return EBI_INVALID_ARGUMENT;
  std::vector<ExprProgram> progs(n_aggs);
  std::vector<int> only(n_aggs), used(n_cols, 0);
  bool any_expr = false;
  for (size_t a = 0; a < n_aggs; ++a) {
    std::vector<int> u(n_cols, 0);
    int rc = check_program(aggs[a], n_cols, &progs[a], &only[a], u);
    if (rc) return fail(ctx, rc, "aggregate: invalid expression");
    if (aggs[a].kind < EBI_AGG_NONE || aggs[a].kind > EBI_AGG_SUM) return EBI_INVALID_ARGUMENT;
    if (only[a] >= 0 && aggs[a].kind == EBI_AGG_NONE)  // AggregationKind::dispatch: todo!()
      return fail(ctx, EBI_UNSUPPORTED, "aggregate: a bare column without an aggregation");
    if ((aggs[a].kind == EBI_AGG_NONE && (!aggs[a].n_out || (!aggs[a].out_values && aggs[a].capacity))) ||
        (aggs[a].kind != EBI_AGG_NONE && !aggs[a].out_scalar))
      return EBI_INVALID_ARGUMENT;
    if (only[a] < 0) {
      any_expr = true;
      for (size_t i = 0; i < n_cols; ++i) used[i] |= u[i];
    }
  }
  // every decoder must hold decoder 0's chunks (api/decoder.rs:902-905)
  const size_t nc = cols[0]->c.chunks.size();
  if (any_expr)
    for (size_t i = 0; i < n_cols; ++i)
      if (cols[i]->c.chunks.size() < nc) return fail(ctx, EBI_PRECONDITION, "aggregate: a column has fewer chunks");
  // bare columns: the decoder's own (in-situ) sum / min / max
  for (size_t a = 0; a < n_aggs; ++a) {
    if (only[a] < 0) continue;
    ebi_query q{};
    q.op = aggs[a].kind == EBI_AGG_SUM ? EBI_OP_SUM : aggs[a].kind == EBI_AGG_MIN ? EBI_OP_MIN : EBI_OP_MAX;
    q.col = cols[only[a]];
    q.bitset = bitset;
    q.chunk_id = -1;
    q.out_scalar = aggs[a].out_scalar;
    const int rc = ebi_hip_query(ctx, &q);
    if (rc) return rc;
  }
  if (!any_expr) return EBI_OK;
  // Each used column: its records decoded, the selected records (bitmask,
  // chunks below nc; the records of each chunk before its NaN end) compacted
  // in order, and their count per chunk -- Expr::Index's materialize of each
  // chunk (decoder/expr.rs:100-121).
  Scratch sc(ctx);
  std::vector<double *> comp(n_cols, nullptr);
  std::vector<uint64_t> col_off((size_t)n_cols * nc, 0), counts_all((size_t)n_cols * nc, 0);
  for (size_t i = 0; i < n_cols; ++i) {
    if (!used[i]) continue;
    Column &c = const_cast<ebi_hip_col *>(cols[i])->c;
    const uint64_t N = c.meta.number_of_records, words = (N + 63) / 64;
    const size_t cn = c.chunks.size();
    double *values = sc.alloc<double>(N);
    uint64_t *mask = bitset ? sc.alloc<uint64_t>(words + 2) : nullptr;
    uint64_t *bits = sc.alloc<uint64_t>(words + 2);
    uint32_t *sel = sc.alloc<uint32_t>(cn);
    uint64_t *counts = sc.alloc<uint64_t>(cn);
    uint64_t *blocks = sc.alloc<uint64_t>(launch::compact_blocks(words) + 2);
    comp[i] = sc.alloc<double>(N);
    if (!values || (bitset && !mask) || !bits || !sel || !counts || !blocks || !comp[i])
      return fail(ctx, EBI_OOM, "aggregate scratch");
    int rc = decode_column(ctx, c, values);
    if (rc) return rc;
    std::vector<uint32_t> hsel(cn);
    for (size_t k = 0; k < cn; ++k) hsel[k] = k < nc ? 1u : 0u;
    hipError_t e = hipMemcpyAsync(sel, hsel.data(), cn * 4, hipMemcpyHostToDevice, ctx->stream);
    if (mask) {
      launch::zero_u32(ctx->stream, (uint32_t *)mask, (uint32_t)((words + 2) * 2));
      if (e == hipSuccess && words) e = hipMemcpyAsync(mask, bitset, words * 8, hipMemcpyHostToDevice, ctx->stream);
    }
    launch::zero_u32(ctx->stream, (uint32_t *)bits, (uint32_t)((words + 2) * 2));
    if ((rc = upload_view(ctx, c, values))) return rc;
    QueryArgs qa{};
    qa.mask = mask;
    qa.out_bits = bits;
    qa.chunk_sel = sel;
    launch::filter(ctx->stream, c.d_view, c.d_rec_jobs, c.n_rec_jobs, qa, 0);
    launch::compact(ctx->stream, bits, words, values, comp[i], blocks, blocks + launch::compact_blocks(words) + 1);
    launch::chunk_count(ctx->stream, c.d_chunks, (uint32_t)cn, bits, counts);
    std::vector<uint64_t> hc(cn);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && cn) e = hipMemcpyAsync(hc.data(), counts, cn * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "aggregate selection");
    uint64_t acc = 0;
    for (size_t k = 0; k < nc; ++k) {
      col_off[i * nc + k] = acc;
      counts_all[i * nc + k] = hc[k];
      acc += hc[k];
    }
  }
  const double *const *dcols = sc.upload(comp);
  const uint64_t *doff = sc.upload(col_off);
  if (!dcols || !doff) return fail(ctx, EBI_OOM, "aggregate tables");
  for (size_t a = 0; a < n_aggs; ++a) {
    if (only[a] >= 0) continue;
    // expression records of chunk i: the shortest of its columns' selections
    // (the zips of decoder/expr.rs:78-96)
    std::vector<int> u(n_cols, 0);
    for (uint32_t p = 0; p < progs[a].n; ++p)
      if (progs[a].nodes[p].kind == EBI_EXPR_INDEX) u[progs[a].nodes[p].index] = 1;
    std::vector<uint64_t> len(nc), out_off(nc + 1, 0);
    std::vector<uint2> tiles;
    for (size_t k = 0; k < nc; ++k) {
      uint64_t m = UINT64_MAX;
      for (size_t i = 0; i < n_cols; ++i)
        if (u[i]) m = std::min(m, counts_all[i * nc + k]);
      len[k] = m;
      out_off[k + 1] = out_off[k] + m;
      for (uint64_t k0 = 0; k0 < m; k0 += kTile) tiles.push_back(make_uint2((uint32_t)k, (uint32_t)k0));
    }
    const uint64_t n_rec = out_off[nc];
    const int kind = aggs[a].kind;
    if (kind == EBI_AGG_NONE && n_rec > aggs[a].capacity)
      return fail(ctx, EBI_INVALID_ARGUMENT, "aggregate: output capacity below the expression's records");
    const uint64_t *dlen = sc.upload(len);
    const uint64_t *dout_off = sc.upload(out_off);
    const uint2 *dtiles = sc.upload(tiles);
    double *partials = sc.alloc<double>(2 * tiles.size());
    unsigned long long *acc = sc.alloc<unsigned long long>(1);
    double *vals = kind == EBI_AGG_NONE ? sc.alloc<double>(n_rec) : nullptr;
    if (!dlen || !dout_off || !dtiles || !partials || !acc || (kind == EBI_AGG_NONE && !vals))
      return fail(ctx, EBI_OOM, "aggregate scratch");
    // kernel kinds: 0 values, 1 max, 2 min, 3 sum
    const int kk = kind == EBI_AGG_NONE ? 0 : kind == EBI_AGG_MAX ? 1 : kind == EBI_AGG_MIN ? 2 : 3;
    launch::expr(ctx->stream, progs[a], dcols, doff, (uint32_t)nc, dlen, dout_off, dtiles, (uint32_t)tiles.size(), kk,
                 partials, acc, vals);
    hipError_t e = hipGetLastError();
    std::vector<double> hp(2 * tiles.size());
    unsigned long long key = 0;
    if (e == hipSuccess && kk == 3 && !tiles.empty())
      e = hipMemcpyAsync(hp.data(), partials, hp.size() * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && (kk == 1 || kk == 2)) e = hipMemcpyAsync(&key, acc, 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && kk == 0 && n_rec)
      e = hipMemcpyAsync(aggs[a].out_values, vals, n_rec * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "aggregate");
    if (kk == 0) {
      *aggs[a].n_out = (size_t)n_rec;
    } else if (kk == 3) {
      // tiles in chunk order, each a double-double: the total in double-double
      double hi = 0.0, lo = 0.0;
      for (size_t t = 0; t < tiles.size(); ++t) {
        dd_acc(hi, lo, hp[2 * t]);
        dd_acc(hi, lo, hp[2 * t + 1]);
      }
      *aggs[a].out_scalar = std::isfinite(hi) ? hi + lo : hi;
    } else {
      // AggregationResult::zero, then f64::max / f64::min over the records
      const bool none = kk == 1 ? key == 0 : key == ~0ull;
      *aggs[a].out_scalar = none ? (kk == 1 ? -INFINITY : INFINITY) : key_to_f64(key);
    }
  }
  return EBI_OK;
}
