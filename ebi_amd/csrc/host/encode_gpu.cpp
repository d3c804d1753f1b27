// C ABI of the GPU write path: BUFF and DeltaSprintz columns encoded on the
// device from values in HBM (kernels/encode.hip), giving the same EBI1
// bytes as the host encoder (Encoder::encode, api/encoder.rs:140-178, with
// compressor/buff.rs:222-305 and compressor/sprintz.rs:151-205), so that a
// column decoded, changed or appended in HBM can be re-encoded without
// moving its records through the host.  The host derives each chunk's
// header fields, size and offset from the per-chunk statistics the GPU
// reduced, then writes the file header and footer around the bodies.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../../include/ebi_hip.h"
#include "../kernels/launch.hpp"
#include "column.hpp"
#include "internal.hpp"

using namespace ebi;

namespace {

int encode_on_device(ebi_hip_ctx *ctx, const double *d_values, size_t n, const ebi_encoder_config *cfg,
                     uint8_t **out, size_t *out_len) {
  if (!cfg || !out || !out_len || (!d_values && n)) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  *out_len = 0;
  const int scheme = cfg->scheme;
  if (scheme != kBUFF && scheme != kDeltaSprintz)
    return fail(ctx, EBI_UNSUPPORTED, "encode_device: BUFF and DeltaSprintz only (ebi_encode covers the others)");
  uint32_t frac = 0;
  if (scheme == kBUFF && buff_frac_bits(cfg->scale, &frac) != 0) return EBI_INVALID_ARGUMENT;
  if (scheme == kDeltaSprintz && cfg->scale > UINT32_MAX)
    return fail(ctx, EBI_INVALID_ARGUMENT, "encode_device: DeltaSprintz scale above 2^32");
  EncodeGeom g;
  int rc = encode_geometry(*cfg, n, &g);
  if (rc) return rc;
  const int codec = scheme == kBUFF ? kCodecBUFF : kCodecSprintz;
  const double scale = (double)cfg->scale;
  EncStat *d_stats = nullptr;
  EncChunk *d_chunks = nullptr;
  uint8_t *d_body = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(ctx->stream);
    hipFree(d_stats);
    hipFree(d_chunks);
    hipFree(d_body);
  };
  const size_t nc = g.n_chunks;
  hipError_t e = hipMalloc((void **)&d_stats, std::max<size_t>(nc, 1) * sizeof(EncStat));
  if (e != hipSuccess) return hip_fail(ctx, e, "encode_device statistics");
  launch::enc_stats(ctx->stream, d_values, n, g.per_chunk, nc, codec, (int32_t)frac, scale, d_stats);
  std::vector<EncStat> st(nc);
  e = hipGetLastError();
  if (e == hipSuccess && nc) e = hipMemcpyAsync(st.data(), d_stats, nc * sizeof(EncStat), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(ctx, e, "encode_device statistics");
  }
  // each chunk's bit width, body size and place in the file
  std::vector<EncChunk> ch(nc);
  std::vector<uint64_t> phys(nc + 1);
  phys[0] = kFileHeaderSize;
  for (size_t c = 0; c < nc; ++c) {
    const uint64_t cn = std::min<uint64_t>(n, (c + 1) * g.per_chunk) - c * g.per_chunk;
    uint64_t size;
    uint32_t bits;
    if (codec == kCodecBUFF) {
      const uint64_t span = (uint64_t)st[c].b - (uint64_t)st[c].a;
      bits = span ? 64u - (uint32_t)__builtin_clzll(span) : 0u;
      const uint64_t planes = bits >= 8 ? bits / 8 : 0, r = bits >= 8 ? bits % 8 : bits;
      size = 20 + (bits ? planes * cn + (cn * r + 7) / 8 : 0);
      ch[c].base = st[c].a;
    } else {
      const uint64_t orz = (uint64_t)st[c].b;
      bits = orz ? 64u - (uint32_t)__builtin_clzll(orz) : 0u;
      size = (104 + cn * bits + 7) / 8;
      ch[c].base = 0;
    }
    ch[c].bits = bits;
    ch[c].frac = frac;
    ch[c].off = phys[c] - kFileHeaderSize;
    phys[c + 1] = phys[c] + size;
  }
  const uint64_t body = phys[nc] - kFileHeaderSize;
  const size_t total = encoded_file_size(scheme, body, nc);
  uint8_t *buf = (uint8_t *)std::malloc(total);
  if (!buf) {
    cleanup();
    return EBI_OOM;
  }
  e = hipMalloc((void **)&d_chunks, std::max<size_t>(nc, 1) * sizeof(EncChunk));
  if (e == hipSuccess) e = hipMalloc((void **)&d_body, (body + 7) / 4 * 4 + 4);
  if (e == hipSuccess && nc)
    e = hipMemcpyAsync(d_chunks, ch.data(), nc * sizeof(EncChunk), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    launch::zero_u32(ctx->stream, (uint32_t *)d_body, (uint32_t)((body + 7) / 4));
    launch::enc_pack(ctx->stream, d_values, n, g.per_chunk, nc, codec, d_chunks, scale, cfg->scale, d_body);
    e = hipGetLastError();
  }
  if (e == hipSuccess && body)
    e = hipMemcpyAsync(buf + kFileHeaderSize, d_body, body, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  cleanup();
  if (e != hipSuccess) {
    std::free(buf);
    return hip_fail(ctx, e, "encode_device");
  }
  write_header_footer(buf, *cfg, n, g, phys);
  *out = buf;
  *out_len = total;
  return EBI_OK;
}

}  // namespace

extern "C" int ebi_hip_encode_device(ebi_hip_ctx *ctx, const double *d_values, size_t n,
                                     const ebi_encoder_config *cfg, uint8_t **out, size_t *out_len) {
  EBI_GUARD(ctx);
  return encode_on_device(ctx, d_values, n, cfg, out, out_len);
}

extern "C" int ebi_hip_column_encode(ebi_hip_ctx *ctx, const double *d_values, size_t n,
                                     const ebi_encoder_config *cfg, ebi_hip_col **out) {
  EBI_GUARD(ctx);
  if (!out) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  uint8_t *file = nullptr;
  size_t len = 0;
  int rc = encode_on_device(ctx, d_values, n, cfg, &file, &len);
  if (rc) return rc;
  rc = ebi_hip_column_load(ctx, file, len, out);
  std::free(file);
  return rc;
}

extern "C" int ebi_hip_copy_to_device(ebi_hip_ctx *ctx, void *d_dst, const void *h_src, size_t bytes) {
  EBI_GUARD(ctx);
  if (!bytes) return EBI_OK;
  if (!d_dst || !h_src) return EBI_INVALID_ARGUMENT;
  hipError_t e = hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "copy to device");
  return EBI_OK;
}
