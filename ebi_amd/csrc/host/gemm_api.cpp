// C ABI of the decompress-then-matmul path: decode resident columns into HBM,
// then multiply on the FP64 matrix cores (kernels/dgemm.hip).
//
//  * ebi_hip_matmul mirrors Decoder::matmul_cuda (lemolatoon/ebi
//    ebi/src/api/decoder.rs:659-724): the column holds batch = N / (m * k)
//    row-major data matrices A_b (m x k) back to back; the host passes the
//    row-major target B (k x n); the result is batch row-major C_b = A_b * B.
//  * ebi_hip_gemm_decoded multiplies two whole columns: A (m x k) from one
//    column and B (k x n) from another, e.g. matrices stored row-chunked
//    (one row per chunk), which the reference's one-matrix-per-chunk
//    requirement (ebi/src/decoder/chunk_reader.rs:261-280) cannot hold.
#include <cstring>

#include "../../../include/ebi_hip.h"
#include "../kernels/launch.hpp"
#include "column.hpp"
#include "internal.hpp"

namespace ebi {
namespace {

// Decodes every chunk of `c` into device buffer `out` (N doubles).
int decode_into(ebi_hip_ctx *ctx, Column &c, double *out) {
  int rc = check_chunk_errors(ctx, c, -1);
  if (rc) return rc;
  if ((rc = upload_view(ctx, c, out))) return rc;
  return enqueue_decode(ctx, c);
}

void record_phases(ebi_hip_ctx *ctx, int n_phases) {
  double total = 0.0;
  for (int p = 0; p < 4; ++p) {
    float ms = 0.0f;
    if (p < n_phases) hipEventElapsedTime(&ms, ctx->ev_ph[p], ctx->ev_ph[p + 1]);  // n_phases <= 3
    ctx->phase_ms[p] = p < n_phases ? ms : 0.0;
    total += ctx->phase_ms[p];
  }
  ctx->last_ms = total;
}

bool mul_overflows(uint64_t a, uint64_t b, uint64_t *out) {
  if (a && b > UINT64_MAX / a) return true;
  *out = a * b;
  return false;
}

}  // namespace
}  // namespace ebi

using namespace ebi;

extern "C" int ebi_hip_matmul(ebi_hip_ctx *ctx, const ebi_hip_col *colc, const double *target, size_t target_rows,
                              size_t target_cols, size_t data_rows, size_t data_cols, double *out) {
  EBI_GUARD(ctx);
  if (!colc || !target || !out) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  const uint64_t N = c.meta.number_of_records;
  uint64_t data_size = 0, target_size = 0, out_size = 0;
  // Preconditions of matmul_cuda (api/decoder.rs:672-684) and the positive
  // m, n, k size check of cuda_binding::Context::gemm (lib.rs:66-72).
  if (data_cols != target_rows || !data_rows || !data_cols || !target_cols ||
      mul_overflows(data_rows, data_cols, &data_size) || mul_overflows(target_rows, target_cols, &target_size) ||
      N % data_size != 0)
    return fail(ctx, EBI_PRECONDITION, "matmul: matrix shapes do not match the column");
  // a chunk that ends at a NaN holds fewer records than its data matrices
  // (the reference's slice of the decoded chunk, chunk_reader.rs:271, panics)
  if (c.truncated) return fail(ctx, EBI_PRECONDITION, "matmul: a chunk ends at a NaN before its record count");
  if (data_rows > UINT32_MAX || data_cols > UINT32_MAX || target_cols > UINT32_MAX)
    return fail(ctx, EBI_UNSUPPORTED, "matmul: matrix dimension above 2^32");
  const uint64_t batch = N / data_size;
  if (mul_overflows(batch, (uint64_t)data_rows * target_cols, &out_size))
    return fail(ctx, EBI_UNSUPPORTED, "matmul: result too large");
  if (batch == 0) return EBI_OK;
  if (batch >= (1ull << 32)) return fail(ctx, EBI_UNSUPPORTED, "matmul: more than 2^32 data matrices");
  if (ctx->gemm_a.ensure(N * 8) || ctx->gemm_b.ensure(target_size * 8) || ctx->gemm_c.ensure(out_size * 8))
    return fail(ctx, EBI_OOM, "matmul scratch allocation");
  double *dA = (double *)ctx->gemm_a.p, *dB = (double *)ctx->gemm_b.p, *dC = (double *)ctx->gemm_c.p;
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  hipEventRecord(ctx->ev_ph[0], ctx->stream);
  int rc = decode_into(ctx, c, dA);
  if (rc) return rc;
  hipEventRecord(ctx->ev_ph[1], ctx->stream);
  hipError_t e = hipMemcpyAsync(dB, target, target_size * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "matmul target upload");
  hipEventRecord(ctx->ev_ph[2], ctx->stream);
  e = launch::dgemm(ctx->stream, dA, dB, dC, (uint32_t)data_rows, (uint32_t)target_cols, (uint32_t)data_cols, batch,
                    data_size, 0, (uint64_t)data_rows * target_cols);
  if (e != hipSuccess) return hip_fail(ctx, e, "dgemm launch");
  hipEventRecord(ctx->ev_ph[3], ctx->stream);
  if ((rc = sync_and_check(ctx, "matmul"))) return rc;
  record_phases(ctx, 3);
  e = hipMemcpy(out, dC, out_size * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(ctx, e, "matmul result copy");
  return EBI_OK;
}

extern "C" int ebi_hip_gemm_decoded(ebi_hip_ctx *ctx, const ebi_hip_col *col_a, const ebi_hip_col *col_b, size_t m,
                                    size_t n, size_t k, double *c_host, double *c_device) {
  EBI_GUARD(ctx);
  if (!col_a || !col_b || (!c_host && !c_device)) return EBI_INVALID_ARGUMENT;
  Column &a = const_cast<ebi_hip_col *>(col_a)->c;
  Column &b = const_cast<ebi_hip_col *>(col_b)->c;
  uint64_t mk = 0, kn = 0, mn = 0;
  if (!m || !n || !k || mul_overflows(m, k, &mk) || mul_overflows(k, n, &kn) || mul_overflows(m, n, &mn) ||
      a.meta.number_of_records != mk || b.meta.number_of_records != kn)
    return fail(ctx, EBI_PRECONDITION, "gemm_decoded: column sizes are not m*k and k*n");
  if (a.truncated || b.truncated)
    return fail(ctx, EBI_PRECONDITION, "gemm_decoded: a chunk ends at a NaN before its record count");
  if (m > UINT32_MAX || n > UINT32_MAX || k > UINT32_MAX)
    return fail(ctx, EBI_UNSUPPORTED, "gemm_decoded: matrix dimension above 2^32");
  if (ctx->gemm_a.ensure(mk * 8) || ctx->gemm_b.ensure(kn * 8) || (!c_device && ctx->gemm_c.ensure(mn * 8)))
    return fail(ctx, EBI_OOM, "gemm scratch allocation");
  double *dA = (double *)ctx->gemm_a.p, *dB = (double *)ctx->gemm_b.p;
  double *dC = c_device ? c_device : (double *)ctx->gemm_c.p;
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  hipEventRecord(ctx->ev_ph[0], ctx->stream);
  int rc = decode_into(ctx, a, dA);
  if (rc) return rc;
  hipEventRecord(ctx->ev_ph[1], ctx->stream);
  if ((rc = decode_into(ctx, b, dB))) return rc;
  hipEventRecord(ctx->ev_ph[2], ctx->stream);
  hipError_t e = launch::dgemm(ctx->stream, dA, dB, dC, (uint32_t)m, (uint32_t)n, (uint32_t)k, 1, 0, 0, 0);
  if (e != hipSuccess) return hip_fail(ctx, e, "dgemm launch");
  hipEventRecord(ctx->ev_ph[3], ctx->stream);
  if ((rc = sync_and_check(ctx, "gemm_decoded"))) return rc;
  record_phases(ctx, 3);
  if (c_host) {
    e = hipMemcpy(c_host, dC, mn * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "gemm result copy");
  }
  return EBI_OK;
}

extern "C" int ebi_hip_dgemm_device(ebi_hip_ctx *ctx, const double *d_a, const double *d_b, double *d_c, size_t m,
                                    size_t n, size_t k) {
  EBI_GUARD(ctx);
  if (!d_a || !d_b || !d_c || !m || !n || !k) return EBI_INVALID_ARGUMENT;
  if (m > UINT32_MAX || n > UINT32_MAX || k > UINT32_MAX) return fail(ctx, EBI_UNSUPPORTED, "dgemm: dimension above 2^32");
  hipEventRecord(ctx->ev_ph[0], ctx->stream);
  hipError_t e = launch::dgemm(ctx->stream, d_a, d_b, d_c, (uint32_t)m, (uint32_t)n, (uint32_t)k, 1, 0, 0, 0);
  if (e != hipSuccess) return hip_fail(ctx, e, "dgemm launch");
  hipEventRecord(ctx->ev_ph[1], ctx->stream);
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "dgemm");
  record_phases(ctx, 1);
  return EBI_OK;
}
