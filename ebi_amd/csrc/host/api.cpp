// C ABI of libebi_hip: context, resident columns and the decode entry points.
// Query entry points live in query_api.cpp.
//
// Boundary citations (lemolatoon/ebi): the context replaces
// cuda_binding::Context (cuda-binding/src/lib.rs:32-181); column_load is
// Decoder::new (api/decoder.rs:136-156) plus the upload the reference never
// does; decode_device is the whole-column materialize (api/decoder.rs:202-237)
// written to HBM.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>

#include "../../../include/ebi_hip.h"
#include "../kernels/launch.hpp"
#include "column.hpp"
#include "internal.hpp"

namespace ebi {

int DevBuf::ensure(size_t bytes) {
  if (bytes <= cap && p) return 0;
  if (p) hipFree(p);
  p = nullptr;
  cap = 0;
  size_t want = std::max<size_t>(bytes, 1 << 20);
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    return (int)e;
  }
  cap = want;
  return 0;
}
void DevBuf::release() {
  if (p) hipFree(p);
  p = nullptr;
  cap = 0;
}

int fail(ebi_hip_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

int hip_fail(ebi_hip_ctx *ctx, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, e == hipErrorOutOfMemory ? EBI_OOM : EBI_HIP, m);
}

int map_dev_err(uint32_t flags) {
  if (!flags) return EBI_OK;
  if (flags & kErrCorrupt) return EBI_CONVERSION;
  return EBI_UNEXPECTED_END;
}

static uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

static uint64_t be64(const uint8_t *p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v = (v << 8) | p[i];
  return v;
}
static uint32_t be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

bool is_xor_codec(int scheme) {
  return scheme == kGorilla || scheme == kChimp || scheme == kElf || scheme == kElfOnChimp;
}

std::vector<Job> tile_jobs_of(const Column &c, uint32_t per_job, std::vector<uint32_t> *first) {
  std::vector<Job> jobs;
  const size_t nc = c.chunks.size();
  if (first) first->assign(nc + 1, 0);
  for (size_t k = 0; k < nc; ++k) {
    if (first) (*first)[k] = (uint32_t)jobs.size();
    const HostChunk &h = c.chunks[k];
    if (c.xor_codec || c.chimp128) continue;
    if (c.meta.scheme == kRLE) {
      if (h.err) continue;  // never decoded
      for (uint32_t t = 0; t < h.n_tiles; ++t) {
        const uint32_t steps = c.rle_steps[h.dev.tile0 + t];
        for (uint32_t s0 = 0; s0 < steps; s0 += per_job)
          jobs.push_back(Job{std::min(per_job, steps - s0), 0, (uint32_t)k, t | (s0 << 24)});
      }
    } else {
      for (uint32_t t = 0; t < h.n_tiles; t += per_job)
        jobs.push_back(Job{std::min(per_job, h.n_tiles - t), 0, (uint32_t)k, t});
    }
  }
  if (first) (*first)[nc] = (uint32_t)jobs.size();
  return jobs;
}

ColView view_of(const Column &c, double *out) {
  ColView v{};
  v.chunks = c.d_chunks;
  v.tile_index = c.d_tile_index;
  v.out = out;
  v.chunk_err = c.d_chunk_err;
  v.tile_jobs = c.d_tile_jobs;
  v.seg_jobs = c.d_seg_jobs;
  v.serial_jobs = c.d_serial_jobs;
  v.n_records = c.meta.number_of_records;
  v.n_chunks = (uint32_t)c.chunks.size();
  v.n_tile_jobs = c.n_tile_jobs;
  v.n_seg_jobs = c.n_seg_jobs;
  v.n_serial_jobs = c.n_serial_jobs;
  v.ilv = c.d_ilv;
  v.c128_ring = c.d_c128_ring;
  v.chunk_end = c.d_chunk_end;
  v.precision = c.precision;
  v.codec = (uint32_t)c.meta.scheme;
  return v;
}

int upload_view(ebi_hip_ctx *ctx, Column &c, double *out) {
  ColView v = view_of(c, out);
  *ctx->h_view = v;
  hipError_t e = hipMemcpyAsync(c.d_view, ctx->h_view, sizeof(ColView), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpyAsync(ColView)");
  // The pinned staging slot is reused by the next call: wait for the copy.
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  return EBI_OK;
}

hipError_t stream_wait(hipStream_t s) {
  hipError_t e;
  while ((e = hipStreamQuery(s)) == hipErrorNotReady) {
  }
  return e;
}

int sync_and_check(ebi_hip_ctx *ctx, const char *what) {
  hipError_t e = hipMemcpyAsync(ctx->h_flags, ctx->d_flags, 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = stream_wait(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  uint32_t f = ctx->h_flags[0];
  if (f) {
    char buf[128];
    std::snprintf(buf, sizeof buf, "%s: device reported chunk error flags 0x%x", what, f);
    return fail(ctx, map_dev_err(f), buf);
  }
  return EBI_OK;
}

int check_chunk_errors(ebi_hip_ctx *ctx, const Column &c, int64_t chunk_id) {
  for (size_t k = 0; k < c.chunks.size(); ++k) {
    if (chunk_id >= 0 && (uint64_t)chunk_id != k) continue;
    if (c.chunks[k].err) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "chunk %zu failed validation at load (flags 0x%x)", k, c.chunks[k].err);
      return fail(ctx, map_dev_err(c.chunks[k].err), buf);
    }
  }
  return EBI_OK;
}

// Queue the kernels that write every record of the selected chunks of `c`
// to out (ColView already uploaded).  chunk_id < 0: all chunks.
int enqueue_decode(ebi_hip_ctx *ctx, Column &c) {
  // Chimp128 always decodes in segment jobs (parse and values in one pass)
  const uint32_t segs = ctx->xor_mode == 0 || c.chimp128 ? c.n_seg_jobs : 0;
  launch::decode(ctx->stream, c.d_view, 1, c.d_prefix, c.n_tile_jobs, c.d_prefix + 6, segs, ctx->d_flags,
                 (uint32_t)ctx->cu_count * ctx->blocks_per_cu, ctx->ev_seg, false, 1u << c.meta.scheme);
  hipEventRecord(ctx->ev_mid, ctx->stream);
  if (c.xor_codec && ctx->xor_mode != 0)
    launch::serial(ctx->stream, c.d_view, c.d_prefix + 2, 1, c.n_serial_jobs, nullptr, ctx->d_flags, true, false);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, "decode launch");
  return EBI_OK;
}

static void free_column(Column &c) {
  hipFree(c.d_body);
  hipFree(c.d_chunks);
  hipFree(c.d_tile_index);
  hipFree(c.d_ckpt);
  hipFree(c.d_ilv);
  hipFree(c.d_c128_ring);
  hipFree(c.d_chunk_err);
  hipFree(c.d_chunk_end);
  hipFree(c.d_view);
  hipFree(c.d_prefix);
  hipFree(c.d_tile_jobs);
  hipFree(c.d_serial_jobs);
  hipFree(c.d_seg_jobs);
  hipFree(c.d_single_jobs);
  hipFree(c.d_rec_jobs);
  hipFree(c.d_span_jobs);
  hipFree(c.d_field_spans);
}

// Chunks that end at a NaN before their footer record count (the load-time
// passes wrote the NaN's record index to d_chunk_end, which starts as the
// footer counts): their descriptors get the shortened count, on the host
// and in the device table.  The reference's XOR readers stop at the first
// NaN and return the records before it without an error
// (chunk_reader/gorilla.rs:149-152, chimp.rs:119, chimp_n.rs:219-265,
// elf.rs:105-258), so every later query sees the shortened chunk.
static int apply_chunk_ends(ebi_hip_ctx *ctx, Column &col) {
  const size_t nc = col.chunks.size();
  if (!nc || !col.d_chunk_end) return EBI_OK;
  std::vector<uint32_t> ends(nc);
  hipError_t e = hipMemcpy(ends.data(), col.d_chunk_end, nc * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(ctx, e, "chunk ends");
  bool changed = false;
  for (size_t k = 0; k < nc; ++k) {
    DevChunk &d = col.chunks[k].dev;
    if (ends[k] < d.n && !col.chunks[k].err) {
      d.n = ends[k];
      changed = true;
    }
  }
  if (!changed) return EBI_OK;
  col.truncated = true;
  std::vector<DevChunk> table(nc);
  for (size_t k = 0; k < nc; ++k) table[k] = col.chunks[k].dev;
  e = hipMemcpy(col.d_chunks, table.data(), nc * sizeof(DevChunk), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(ctx, e, "chunk table (NaN ends)");
  return EBI_OK;
}

template <class T>
static hipError_t upload_vec(T **dst, const std::vector<T> &v) {
  *dst = nullptr;
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void **)dst, v.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// Parallel checkpoint index of an XOR column (Gorilla, Chimp, Chimp128, Elf,
// ElfOnChimp; kernels/xindex.hpp): candidates over spans of every chunk's
// stream (Gorilla: after its walkers have parsed the first kHintSpans
// spans), the per-chunk walkers, then the checkpoints and per-chunk
// events.  The view must be uploaded; errors land in ctx->d_flags[0] and
// d_chunk_err, NaN ends in d_chunk_end, as with the lane-per-chunk pass.
static int build_xindex(ebi_hip_ctx *ctx, Column &col) {
  const size_t nc = col.chunks.size();
  std::vector<uint32_t> first(nc + 1, 0);
  std::vector<uint2> groups;
  for (size_t k = 0; k < nc; ++k) {
    const uint32_t ns = span_count(col.chunks[k].dev.size * 8u);
    for (uint32_t j = 0; j < ns; j += kSpansPerLane) groups.push_back(make_uint2((uint32_t)k, j));
    if ((uint64_t)first[k] + ns > UINT32_MAX) return fail(ctx, EBI_OOM, "XOR index: too many spans");
    first[k + 1] = first[k] + ns;
  }
  const uint64_t spans = first[nc];
  if (groups.size() > UINT32_MAX) return fail(ctx, EBI_OOM, "XOR index: too many span groups");
  XIndexArgs a{};
  a.cols = col.d_view;
  a.err = ctx->d_flags;
  a.nc = (uint32_t)nc;
  a.n_groups = (uint32_t)groups.size();
  // one allocation: span_first, groups, cand, res, pick, walk, ev
  const size_t o_first = 0, o_groups = round_up((nc + 1) * 4, 256),
               o_cand = o_groups + round_up(groups.size() * 8, 256), o_res = o_cand + round_up(spans * 8, 256),
               o_pick = o_res + round_up(spans * sizeof(SpanSum), 256),
               o_walk = o_pick + round_up(spans * sizeof(SpanPick), 256),
               o_ev = o_walk + round_up(nc * sizeof(ChunkWalk), 256), total = o_ev + round_up(nc * 8, 256) + 256;
  uint8_t *d = nullptr;
  hipError_t e = hipMalloc((void **)&d, total);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc(XOR index scratch)");
  a.span_first = reinterpret_cast<uint32_t *>(d + o_first);
  a.groups = reinterpret_cast<uint2 *>(d + o_groups);
  a.cand = reinterpret_cast<uint64_t *>(d + o_cand);
  a.res = reinterpret_cast<SpanSum *>(d + o_res);
  a.pick = reinterpret_cast<SpanPick *>(d + o_pick);
  a.walk = reinterpret_cast<ChunkWalk *>(d + o_walk);
  a.ev = reinterpret_cast<uint64_t *>(d + o_ev);
  e = hipMemcpyAsync(d + o_first, first.data(), (nc + 1) * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && !groups.empty())
    e = hipMemcpyAsync(d + o_groups, groups.data(), groups.size() * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    launch::xindex_init(ctx->stream, a, (uint32_t)spans);
    const bool gorilla = nc && col.chunks[0].dev.codec == kCodecGorilla;
    if (gorilla) launch::xindex_walk(ctx->stream, a, kHintSpans, false);  // the state its candidates start in
    launch::xindex_cands(ctx->stream, a);
    launch::xindex_walk(ctx->stream, a, UINT32_MAX, gorilla);
    launch::xindex_emit(ctx->stream, a);
    e = hipGetLastError();
  }
  const hipError_t e2 = hipStreamSynchronize(ctx->stream);
  hipFree(d);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return hip_fail(ctx, e, "XOR index");
  return EBI_OK;
}

// Lane-interleaved copy of an XOR column's segment streams (after the index
// pass, whose checkpoints give each segment's bits): the row count of every
// segment job, their prefix as each job's first row (Job::col), then the
// copy.  Chunks rejected by the index pass get no rows.
static int build_ilv(ebi_hip_ctx *ctx, Column &col) {
  const uint32_t nj = col.n_seg_jobs;
  uint32_t *d_rows = nullptr;
  hipError_t e = hipMalloc((void **)&d_rows, (size_t)nj * 4);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc(ilv row counts)");
  std::vector<uint32_t> rows(nj);
  launch::ilv_rows(ctx->stream, col.d_view, nj, d_rows);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(rows.data(), d_rows, (size_t)nj * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  uint64_t total = 0;
  for (uint32_t j = 0; j < nj && e == hipSuccess; ++j) {
    col.h_seg_jobs[j].col = (uint32_t)total;
    total += rows[j];
  }
  if (e == hipSuccess && total > UINT32_MAX) {
    hipFree(d_rows);
    return fail(ctx, EBI_OOM, "interleaved segment streams exceed 2^32 rows");
  }
  if (e == hipSuccess) e = hipMalloc((void **)&col.d_ilv, std::max<uint64_t>(total, 1) * kIlvRowBlocks * 16);
  if (e == hipSuccess) e = hipMemcpy(col.d_seg_jobs, col.h_seg_jobs.data(), (size_t)nj * sizeof(Job), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch::ilv_copy(ctx->stream, col.d_view, nj, d_rows, col.d_ilv);
    e = hipGetLastError();
  }
  const hipError_t e2 = hipStreamSynchronize(ctx->stream);
  hipFree(d_rows);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return hip_fail(ctx, e, "interleaved segment streams");
  col.ilv_rows = total;
  col.device_bytes += total * kIlvRowBlocks * 16;
  return EBI_OK;
}

// Ring checkpoints of a Chimp128 column (after the index pass): each
// segment job keeps the values of the 128 records before it (128 * 8 bytes
// per 1024 records).  Built as a wavefront over the jobs of every chunk:
// step s decodes job s of each chunk (one wave per chunk) from the ring
// step s - 1 wrote and writes the ring of job s + 1, so the column is
// decoded once, in max(jobs per chunk) launches of up to one wave per chunk.
// The decode validates the values too (the parse-only index cannot see a
// wrong terminator) and lowers each chunk's end to its first NaN.
static int build_c128_ring(ebi_hip_ctx *ctx, Column &col) {
  const size_t nc = col.chunks.size();
  std::vector<uint32_t> first(nc + 1, 0);
  for (const Job &jb : col.h_seg_jobs) ++first[jb.chunk + 1];
  uint32_t steps = 0;
  for (size_t k = 0; k < nc; ++k) {
    steps = std::max(steps, first[k + 1]);
    first[k + 1] += first[k];
  }
  uint32_t *d_first = nullptr;
  hipError_t e = hipMalloc((void **)&col.d_c128_ring, (size_t)col.n_seg_jobs * 128 * 8);
  if (e == hipSuccess) e = upload_vec(&d_first, first);
  if (e != hipSuccess) {
    hipFree(d_first);
    return hip_fail(ctx, e, "Chimp128 ring checkpoints: allocation");
  }
  int rc = upload_view(ctx, col, nullptr);
  std::vector<uint32_t> errs(nc);
  if (!rc) {
    for (uint32_t s = 0; s < steps; ++s)
      launch::c128_ring_step(ctx->stream, col.d_view, d_first, (uint32_t)nc, s, col.d_c128_ring, ctx->d_flags);
    e = hipGetLastError();
    if (e == hipSuccess && nc)
      e = hipMemcpyAsync(errs.data(), col.d_chunk_err, nc * 4, hipMemcpyDeviceToHost, ctx->stream);
    const hipError_t e2 = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) rc = hip_fail(ctx, e, "Chimp128 ring checkpoints");
  }
  if (!rc)
    for (size_t k = 0; k < nc; ++k) col.chunks[k].err |= errs[k];
  hipFree(d_first);
  if (!rc) col.device_bytes += (uint64_t)col.n_seg_jobs * 128 * 8;
  return rc;
}

// Parse, validate codec headers, repack into the device layout, upload and
// build the tile / checkpoint index.
static double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

// A piece of a column's device image taken from the file: kPieceCopy
// copies len bytes; kPieceRuns gathers len run bytes src[9 i]; kPieceValues
// gathers len f64 values src[9 i .. 9 i + 8) (RLE's 9-byte fields).
enum : uint8_t { kPieceCopy = 0, kPieceRuns = 1, kPieceValues = 2 };
struct Piece {
  uint64_t dst;
  const uint8_t *src;
  uint64_t len;
  uint8_t kind;
  uint64_t end() const { return dst + (kind == kPieceValues ? 8 * len : len); }
};

constexpr uint64_t kStageBytes = 32ull << 20;  // one half of the upload double buffer
constexpr uint64_t kStageSplit = 4096;         // repack threads split a window at multiples of this

// Bytes [t0, t1) of the image into dst (= image byte t0): the pieces'
// bytes, zeros between them.  Piece boundaries and t0 / t1 are multiples of
// 8 where they cut a kPieceValues piece (pieces start kAlign-aligned and
// windows are cut at multiples of kStageSplit).
static void fill_image(const std::vector<Piece> &pieces, uint64_t t0, uint64_t t1, uint8_t *dst) {
  size_t i = std::upper_bound(pieces.begin(), pieces.end(), t0,
                              [](uint64_t v, const Piece &p) { return v < p.end(); }) - pieces.begin();
  uint64_t at = t0;
  for (; i < pieces.size() && pieces[i].dst < t1; ++i) {
    const Piece &p = pieces[i];
    const uint64_t a = std::max(t0, p.dst), b = std::min(t1, p.end());
    if (a > at) std::memset(dst + (at - t0), 0, a - at);
    uint8_t *o = dst + (a - t0);
    if (p.kind == kPieceCopy) {
      std::memcpy(o, p.src + (a - p.dst), b - a);
    } else if (p.kind == kPieceRuns) {
      for (uint64_t j = a - p.dst; j < b - p.dst; ++j) *o++ = p.src[9 * j];
    } else {
      for (uint64_t j = (a - p.dst) / 8; j < (b - p.dst) / 8; ++j, o += 8) std::memcpy(o, p.src + 9 * j, 8);
    }
    at = b;
  }
  if (t1 > at) std::memset(dst + (at - t0), 0, t1 - at);
}

// Build the image [0, bytes) of `pieces` in device memory d_dst on the
// context's stream: windows of kStageBytes, alternating between the two
// pinned buffers, each repacked on up to 8 host threads and copied with
// hipMemcpyAsync; a buffer is refilled once its previous copy is done.
// Returns with the copies queued (later work on the stream is ordered
// after them).
static int upload_image(ebi_hip_ctx *ctx, const std::vector<Piece> &pieces, uint8_t *d_dst, uint64_t bytes) {
  hipError_t e = hipSuccess;
  for (int s = 0; s < 2 && e == hipSuccess; ++s) {
    if (!ctx->stage[s]) e = hipHostMalloc((void **)&ctx->stage[s], kStageBytes, hipHostMallocDefault);
    if (e == hipSuccess && !ctx->stage_ev[s]) e = hipEventCreateWithFlags(&ctx->stage_ev[s], hipEventDisableTiming);
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "column upload buffers");
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  for (uint64_t w0 = 0, w = 0; w0 < bytes; w0 += kStageBytes, ++w) {
    const int s = (int)(w & 1);
    const uint64_t w1 = std::min(bytes, w0 + kStageBytes);
    // the buffer's last copy (of this or an earlier upload) has ended; an
    // event never recorded counts as complete
    if ((e = hipEventSynchronize(ctx->stage_ev[s])) != hipSuccess) break;
    uint8_t *buf = ctx->stage[s];
    const uint64_t parts = std::min<uint64_t>(std::min(8u, hw), (w1 - w0 + (1u << 20) - 1) >> 20);
    if (parts <= 1) {
      fill_image(pieces, w0, w1, buf);
    } else {
      const uint64_t step = (((w1 - w0) / parts) + kStageSplit - 1) / kStageSplit * kStageSplit;
      std::vector<std::thread> pool;
      for (uint64_t t0 = w0; t0 < w1; t0 += step) {
        const uint64_t t1 = std::min(w1, t0 + step);
        pool.emplace_back([&pieces, t0, t1, buf, w0] { fill_image(pieces, t0, t1, buf + (t0 - w0)); });
      }
      for (std::thread &t : pool) t.join();
    }
    if ((e = hipMemcpyAsync(d_dst + w0, buf, w1 - w0, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess ||
        (e = hipEventRecord(ctx->stage_ev[s], ctx->stream)) != hipSuccess)
      break;
  }
  if (e != hipSuccess) {
    hipStreamSynchronize(ctx->stream);
    return hip_fail(ctx, e, "column upload");
  }
  return EBI_OK;
}

static int load_column(ebi_hip_ctx *ctx, const uint8_t *file, size_t len, Column &col) {
  const auto t0 = std::chrono::steady_clock::now();
  std::string err;
  int rc = parse_file(file, len, &col.meta, &err);
  if (rc) return fail(ctx, rc, err);
  const int scheme = col.meta.scheme;
  if (!(scheme == kUncompressed || scheme == kRLE || scheme == kBUFF || scheme == kGorilla ||
        scheme == kChimp || scheme == kChimp128 || scheme == kElfOnChimp || scheme == kElf ||
        scheme == kDeltaSprintz))
    return fail(ctx, EBI_UNSUPPORTED, "compression scheme outside the accelerated set");
  col.xor_codec = is_xor_codec(scheme);
  col.chimp128 = scheme == kChimp128;
  const size_t nc = col.meta.chunks.size();
  col.chunks.resize(nc);
  struct Body { const uint8_t *src; uint64_t off, len; };
  std::vector<Body> bodies(nc);
  uint64_t dev_bytes = 0, tiles = 0, ckpts = 0;
  for (size_t k = 0; k < nc; ++k) {
    const ChunkGeom &g = col.meta.chunks[k];
    HostChunk &h = col.chunks[k];
    h.geom = g;
    const uint8_t *p = file + g.physical_offset;
    if (g.size >= (1ull << 29) || g.number_of_records >= (1ull << 31))
      return fail(ctx, EBI_UNSUPPORTED, "chunk larger than 512 MiB or 2^31 records");
    const uint32_t n = (uint32_t)g.number_of_records;
    DevChunk d{};
    d.rec_off = g.logical_offset;
    d.n = n;
    d.size = (uint32_t)g.size;
    d.codec = (uint32_t)scheme;
    col.compressed_bytes += g.size;
    uint64_t body_len = g.size;
    const uint8_t *src = p;
    switch (scheme) {
      case kUncompressed:
        if ((uint64_t)n * 8 > g.size) return fail(ctx, EBI_UNEXPECTED_END, "uncompressed chunk shorter than its records");
        body_len = (uint64_t)n * 8;
        h.n_tiles = (n + kTile - 1) / kTile;
        break;
      case kRLE: {
        if (g.size < 8) return fail(ctx, EBI_UNEXPECTED_END, "RLE chunk without field count");
        const uint64_t f = rd_u64(p);
        if (f > (g.size - 8) / 9) return fail(ctx, EBI_UNEXPECTED_END, "RLE fields run past the chunk");
        d.fields = f;
        d.values_off = round_up(f, kAlign);
        body_len = d.values_off + 8 * f;
        h.n_tiles = (uint32_t)((f + kTile - 1) / kTile);
        break;
      }
      case kBUFF: {
        if (g.size < 20) return fail(ctx, EBI_CONVERSION, "BUFF chunk shorter than its header");
        const uint64_t lo = rd_u32(p), hi = rd_u32(p + 4);
        d.base = (int64_t)(lo | (hi << 32));
        const uint32_t nh = rd_u32(p + 8);
        d.bits = rd_u32(p + 12);
        d.frac = rd_u32(p + 16);
        if (nh != n) return fail(ctx, EBI_CONVERSION, "BUFF header record count differs from the footer");
        if (d.bits > 64) return fail(ctx, EBI_CONVERSION, "BUFF bit width above 64");
        if (d.frac > 1074) return fail(ctx, EBI_CONVERSION, "BUFF fractional bits out of range");
        d.n_planes = d.bits >= 8 ? d.bits / 8 : 0;
        d.r = d.bits >= 8 ? d.bits % 8 : d.bits;
        d.plane_stride = (uint32_t)round_up(std::max<uint32_t>(n, 1), kAlign);
        const uint64_t tail = ((uint64_t)n * d.r + 7) / 8;
        if (20 + (uint64_t)d.n_planes * n + tail > g.size)
          return fail(ctx, EBI_CONVERSION, "BUFF planes run past the chunk");
        body_len = (uint64_t)d.n_planes * d.plane_stride + tail;
        h.n_tiles = (n + kTile - 1) / kTile;
        break;
      }
      case kDeltaSprintz: {
        if (g.size < 13) return fail(ctx, EBI_UNEXPECTED_END, "DeltaSprintz chunk shorter than its header");
        d.base = (int64_t)be64(p);
        d.scale = be32(p + 8);
        d.bits = p[12];
        if (d.bits > 64) return fail(ctx, EBI_CONVERSION, "DeltaSprintz bit width above 64");
        if (13 + ((uint64_t)n * d.bits + 7) / 8 > g.size)
          return fail(ctx, EBI_UNEXPECTED_END, "DeltaSprintz fields run past the chunk");
        src = p + 13;
        body_len = g.size - 13;
        h.n_tiles = (n + kTile - 1) / kTile;
        break;
      }
      default:  // XOR codecs: the bit stream as written
        // Chimp128 keeps parse state for records 0..n (the terminator too)
        h.n_segs = col.chimp128 ? n / kSeg128 + 1 : (n + kSeg - 1) / kSeg;
        break;
    }
    bodies[k] = {src, dev_bytes, body_len};
    d.data = dev_bytes;  // relative until the allocation exists
    if (scheme == kDeltaSprintz || scheme == kRLE) {
      d.tile0 = (uint32_t)tiles;
      tiles += h.n_tiles;
    }
    if (h.n_segs) {
      d.xidx = ckpts;  // relative
      ckpts += h.n_segs;
    }
    d.rtile0 = (uint32_t)col.n_rtiles;
    col.n_rtiles += (n + kTile - 1) / kTile;
    h.dev = d;
    dev_bytes += round_up(std::max<uint64_t>(body_len, 1), kAlign) + kPad;
  }
  col.device_bytes = dev_bytes;
  col.tile_index_entries = tiles;
  col.ckpt_entries = ckpts;

  hipError_t e;
  if ((e = hipMalloc((void **)&col.d_body, std::max<uint64_t>(dev_bytes, kAlign))) != hipSuccess)
    return hip_fail(ctx, e, "hipMalloc(column body)");
  // The device image as pieces of the file (everything between them is
  // zero), repacked window by window into the context's pinned double
  // buffer and copied asynchronously, so the copy of one window overlaps the
  // repack of the next.
  {
    std::vector<Piece> pieces;
    pieces.reserve(nc * 2);
    for (size_t k = 0; k < nc; ++k) {
      const Body &b = bodies[k];
      const DevChunk &d = col.chunks[k].dev;
      const uint8_t *p = file + col.chunks[k].geom.physical_offset;
      if (scheme == kRLE) {  // u8 runs, then the f64 values, from 9-byte fields
        if (d.fields) {
          pieces.push_back(Piece{b.off, p + 8, d.fields, kPieceRuns});
          pieces.push_back(Piece{b.off + d.values_off, p + 9, d.fields, kPieceValues});
        }
      } else if (scheme == kBUFF) {  // planes at plane_stride, then the LSB tail
        for (uint32_t pl = 0; pl < d.n_planes; ++pl)
          if (d.n) pieces.push_back(Piece{b.off + (uint64_t)pl * d.plane_stride, p + 20 + (size_t)pl * d.n, d.n, kPieceCopy});
        const uint64_t tail = ((uint64_t)d.n * d.r + 7) / 8;
        if (tail) pieces.push_back(Piece{b.off + (uint64_t)d.n_planes * d.plane_stride, p + 20 + (size_t)d.n_planes * d.n, tail, kPieceCopy});
      } else if (b.len) {
        pieces.push_back(Piece{b.off, b.src, b.len, kPieceCopy});
      }
    }
    if ((rc = upload_image(ctx, pieces, col.d_body, dev_bytes))) return rc;
  }
  if (tiles && (e = hipMalloc((void **)&col.d_tile_index, tiles * 8)) != hipSuccess)
    return hip_fail(ctx, e, "hipMalloc(tile index)");
  if (ckpts && (e = hipMalloc((void **)&col.d_ckpt, ckpts * sizeof(XorCkpt))) != hipSuccess)
    return hip_fail(ctx, e, "hipMalloc(checkpoints)");
  if ((e = hipMalloc((void **)&col.d_chunk_err, std::max<size_t>(nc, 1) * 4)) != hipSuccess)
    return hip_fail(ctx, e, "hipMalloc(chunk errors)");
  if (col.xor_codec || col.chimp128) {
    std::vector<uint32_t> ends(nc);
    for (size_t k = 0; k < nc; ++k) ends[k] = col.chunks[k].dev.n;
    if ((e = upload_vec(&col.d_chunk_end, ends)) != hipSuccess) return hip_fail(ctx, e, "chunk ends");
  }
  if ((e = hipMalloc((void **)&col.d_view, sizeof(ColView))) != hipSuccess)
    return hip_fail(ctx, e, "hipMalloc(view)");

  std::vector<DevChunk> table(nc);
  for (size_t k = 0; k < nc; ++k) {
    DevChunk &d = col.chunks[k].dev;
    d.data = (uint64_t)(uintptr_t)col.d_body + d.data;
    if (col.chunks[k].n_segs) d.xidx = (uint64_t)(uintptr_t)(col.d_ckpt + d.xidx);
    table[k] = d;
  }
  if ((e = upload_vec(&col.d_chunks, table)) != hipSuccess) return hip_fail(ctx, e, "upload chunk table");

  // Resident job lists.  Decode tile jobs cover up to kTilesPerJob
  // consecutive tiles of one chunk (Job::kind = tile count), so a wave pays
  // the descriptor lookups once per up to 16 KiB of output; the load-time
  // index pass walks single tiles.
  std::vector<Job> index_jobs;
  if (scheme != kRLE) col.h_tile_jobs = tile_jobs_of(col, kTilesPerJob, &col.tile_job_first);  // RLE: after the index
  for (size_t k = 0; k < nc; ++k) {
    if (!col.xor_codec && !col.chimp128)
      for (uint32_t t = 0; t < col.chunks[k].n_tiles; ++t) index_jobs.push_back(Job{1, 0, (uint32_t)k, t});
    // Job::col: the job's index (Chimp128's ring checkpoints); the other
    // codecs' first ilv row, set once the ilv is built
    for (uint32_t s = 0; s < col.chunks[k].n_segs; s += kWaveLanes)
      col.h_seg_jobs.push_back(Job{kJobSeg, (uint32_t)col.h_seg_jobs.size(), (uint32_t)k, s});
  }
  if (scheme == kRLE) col.tile_job_first.assign(nc + 1, 0);
  std::vector<Job> rec_jobs, span_jobs, field_spans;
  col.rle_direct = scheme == kRLE;
  // query span length: up to kSpanTiles tiles per wave, fewer when the
  // column would otherwise give under kSpanMinWaves waves (small columns
  // keep the GPU filled; large ones amortize the per-span set-up)
  uint64_t all_tiles = 0, all_fields = 0;
  for (size_t k = 0; k < nc; ++k) {
    all_tiles += (col.chunks[k].dev.n + kTile - 1) / kTile;
    if (scheme == kRLE) all_fields += col.chunks[k].n_tiles;
  }
  auto span_of = [](uint64_t tiles) {
    uint32_t s = kSpanTiles;
    while (s > 1 && tiles / s < kSpanMinWaves) s /= 2;
    return s;
  };
  col.span_tiles = span_of(all_tiles);
  col.field_span_tiles = span_of(all_fields);
  for (size_t k = 0; k < nc; ++k) {
    const uint32_t nt = (col.chunks[k].dev.n + kTile - 1) / kTile, st = col.span_tiles;
    for (uint32_t t = 0; t < nt; ++t) rec_jobs.push_back(Job{kJobTile, 0, (uint32_t)k, t});
    for (uint32_t t = 0; t < nt; t += st) span_jobs.push_back(Job{std::min(st, nt - t), 0, (uint32_t)k, t});
    if (scheme == kRLE) {  // field tiles; their f64 partials use the chunk's record-tile slots
      const uint32_t ft = col.chunks[k].n_tiles, fs = col.field_span_tiles;
      if (col.chunks[k].dev.fields > col.chunks[k].dev.n) col.rle_direct = false;
      for (uint32_t t = 0; t < ft; t += fs) field_spans.push_back(Job{std::min(fs, ft - t), 0, (uint32_t)k, t});
    }
  }
  col.n_rec_jobs = (uint32_t)rec_jobs.size();
  col.n_span_jobs = (uint32_t)span_jobs.size();
  col.n_field_spans = (uint32_t)field_spans.size();
  std::vector<Job> single;
  if (col.xor_codec || col.chimp128) {
    for (size_t k = 0; k < nc; k += kWaveLanes)
      col.h_serial_jobs.push_back(Job{kJobSerial, 0, (uint32_t)k, (uint32_t)std::min<size_t>(kWaveLanes, nc - k)});
    for (size_t k = 0; k < nc; ++k) single.push_back(Job{kJobSerial, 0, (uint32_t)k, 1});
  }
  col.n_tile_jobs = (uint32_t)col.h_tile_jobs.size();
  col.n_serial_jobs = (uint32_t)col.h_serial_jobs.size();
  col.n_seg_jobs = (uint32_t)col.h_seg_jobs.size();
  if ((e = upload_vec(&col.d_tile_jobs, col.h_tile_jobs)) != hipSuccess ||
      (e = upload_vec(&col.d_serial_jobs, col.h_serial_jobs)) != hipSuccess ||
      (e = upload_vec(&col.d_seg_jobs, col.h_seg_jobs)) != hipSuccess ||
      (e = upload_vec(&col.d_single_jobs, single)) != hipSuccess ||
      (e = upload_vec(&col.d_rec_jobs, rec_jobs)) != hipSuccess ||
      (e = upload_vec(&col.d_span_jobs, span_jobs)) != hipSuccess ||
      (e = upload_vec(&col.d_field_spans, field_spans)) != hipSuccess)
    return hip_fail(ctx, e, "upload job lists");

  {
    const std::vector<uint32_t> pre = {0, col.n_tile_jobs, 0, col.n_serial_jobs, 0, (uint32_t)nc, 0, col.n_seg_jobs};
    if ((e = upload_vec(&col.d_prefix, pre)) != hipSuccess) return hip_fail(ctx, e, "upload prefix");
  }
  // Load-time index pass.
  const double upload_ms = ms_since(t0);
  const auto t1 = std::chrono::steady_clock::now();
  launch::zero_u32(ctx->stream, col.d_chunk_err, (uint32_t)nc);
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  if ((rc = upload_view(ctx, col, nullptr))) return rc;
  Job *d_index_jobs = nullptr;
  if (scheme == kDeltaSprintz || scheme == kRLE) {
    if ((e = upload_vec(&d_index_jobs, index_jobs)) != hipSuccess) return hip_fail(ctx, e, "upload index jobs");
    launch::tile_index(ctx->stream, col.d_view, 0, d_index_jobs, (uint32_t)index_jobs.size(), (uint32_t)nc,
                       ctx->d_flags);
  }
  // XOR codecs: the parallel index, or (EBI_HIP_XOR_INDEX=serial) one lane
  // decoding each chunk from its start
  const bool xor_index = col.xor_codec || col.chimp128;
  if (xor_index && ctx->xindex_serial)
    launch::serial(ctx->stream, col.d_view, col.d_prefix + 2, 1, col.n_serial_jobs, nullptr, ctx->d_flags, false,
                   true);
  if ((e = hipGetLastError()) == hipSuccess && xor_index && !ctx->xindex_serial) rc = build_xindex(ctx, col);
  if (e != hipSuccess || rc) {
    hipStreamSynchronize(ctx->stream);
    if (d_index_jobs) hipFree(d_index_jobs);
    return rc ? rc : hip_fail(ctx, e, "index launch");
  }
  std::vector<uint32_t> errs(nc);
  if (nc) e = hipMemcpyAsync(errs.data(), col.d_chunk_err, nc * 4, hipMemcpyDeviceToHost, ctx->stream);
  const hipError_t e2 = hipStreamSynchronize(ctx->stream);
  if (d_index_jobs) hipFree(d_index_jobs);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return hip_fail(ctx, e, "index pass");
  for (size_t k = 0; k < nc; ++k) col.chunks[k].err = errs[k];
  const double index_ms = ms_since(t1);
  const auto t2 = std::chrono::steady_clock::now();
  // Gorilla / Chimp / Elf: the index pass found the NaN ends; Chimp128's
  // parse-only index cannot see values, its ring build's decode finds them
  if (col.xor_codec && (rc = apply_chunk_ends(ctx, col))) return rc;
  if (col.xor_codec && col.n_seg_jobs && (rc = build_ilv(ctx, col))) return rc;
  if (col.chimp128 && col.n_seg_jobs && (rc = build_c128_ring(ctx, col))) return rc;
  if (col.chimp128 && (rc = apply_chunk_ends(ctx, col))) return rc;
  if (scheme == kRLE && nc) {
    // Balance RLE decode jobs by output: a job expands up to kTilesPerJob
    // steps of kTile output records of one field tile (runs of up to 255
    // make a field tile's output 1 to 255 steps).
    uint64_t n_tiles = 0;
    for (size_t k = 0; k < nc; ++k) n_tiles += col.chunks[k].n_tiles;
    std::vector<int64_t> starts(n_tiles);
    if (n_tiles && (e = hipMemcpy(starts.data(), col.d_tile_index, n_tiles * 8, hipMemcpyDeviceToHost)) != hipSuccess)
      return hip_fail(ctx, e, "RLE tile index");
    col.rle_steps.assign(n_tiles, 0);
    for (size_t k = 0; k < nc; ++k) {
      const HostChunk &h = col.chunks[k];
      for (uint32_t t = 0; t < h.n_tiles; ++t) {
        const int64_t a = starts[h.dev.tile0 + t];
        const int64_t b = t + 1 < h.n_tiles ? starts[h.dev.tile0 + t + 1] : (int64_t)h.dev.n;
        col.rle_steps[h.dev.tile0 + t] = b > a ? (uint16_t)((b - a + kTile - 1) / kTile) : 0;
      }
    }
    col.h_tile_jobs = tile_jobs_of(col, kTilesPerJob, &col.tile_job_first);
    col.n_tile_jobs = (uint32_t)col.h_tile_jobs.size();
    hipFree(col.d_tile_jobs);
    col.d_tile_jobs = nullptr;
    const uint32_t jobs = col.n_tile_jobs;
    if ((e = upload_vec(&col.d_tile_jobs, col.h_tile_jobs)) != hipSuccess ||
        (e = hipMemcpy(col.d_prefix + 1, &jobs, 4, hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(ctx, e, "RLE job lists");
  }
  // resident bytes beside the chunk bodies (and the ilv / ring counted as built)
  col.device_bytes += col.ckpt_entries * sizeof(XorCkpt) + col.tile_index_entries * 8 + nc * sizeof(DevChunk) +
                      (uint64_t)(col.h_tile_jobs.size() + col.h_serial_jobs.size() + col.h_seg_jobs.size() +
                                 col.n_rec_jobs + col.n_span_jobs + col.n_field_spans) *
                          sizeof(Job);
  // host clock of the load's phases: parse + repack + upload, the index
  // pass, the lane-interleaved copy / Chimp128 ring / job sizing
  ctx->phase_ms[0] = upload_ms;
  ctx->phase_ms[1] = index_ms;
  ctx->phase_ms[2] = ms_since(t2);
  ctx->phase_ms[3] = 0.0;
  ctx->last_ms = ms_since(t0);
  return EBI_OK;
}

}  // namespace ebi

using namespace ebi;


extern "C" int ebi_hip_ctx_create(int device, ebi_hip_ctx **out) {
  if (!out) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return EBI_HIP;
  if (device < 0 || device >= n) return EBI_INVALID_ARGUMENT;
  if ((e = hipSetDevice(device)) != hipSuccess) return EBI_HIP;
  ebi_hip_ctx *c = new (std::nothrow) ebi_hip_ctx();
  if (!c) return EBI_OOM;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    c->cu_count = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess || hipEventCreate(&c->ev_mid) != hipSuccess ||
      hipEventCreate(&c->ev_seg) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreate(&c->ev_ph[0]) != hipSuccess || hipEventCreate(&c->ev_ph[1]) != hipSuccess ||
      hipEventCreate(&c->ev_ph[2]) != hipSuccess || hipEventCreate(&c->ev_ph[3]) != hipSuccess ||
      hipMalloc((void **)&c->d_flags, 64 * sizeof(uint32_t)) != hipSuccess ||
      hipHostMalloc((void **)&c->h_flags, 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->h_view, 64 * sizeof(ColView), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->h_batch, 1 << 16, hipHostMallocDefault) != hipSuccess) {
    ebi_hip_ctx_destroy(c);
    return EBI_HIP;
  }
  c->h_view_cap = 64;
  c->h_batch_cap = 1 << 16;
  // Powers of ten for the Elf restore, parsed exactly as the reference does.
  static double p10[400], n10[400];
  char buf[32];
  for (int i = 0; i < 400; ++i) {
    std::snprintf(buf, sizeof buf, "1.0E%d", i);
    p10[i] = std::strtod(buf, nullptr);
    std::snprintf(buf, sizeof buf, "1.0E-%d", i);
    n10[i] = std::strtod(buf, nullptr);
  }
  if (launch::set_pow10(p10, n10) != hipSuccess || noalloc_init(c) != EBI_OK) {
    ebi_hip_ctx_destroy(c);
    return EBI_HIP;
  }
  const char *mode = std::getenv("EBI_HIP_XOR_MODE");
  if (mode && std::strcmp(mode, "serial") == 0) c->xor_mode = 1;
  const char *xi = std::getenv("EBI_HIP_XOR_INDEX");
  if (xi && std::strcmp(xi, "serial") == 0) c->xindex_serial = 1;
  // launch-shape overrides (the batch tuner picks its own per batch)
  if (const char *v = std::getenv("EBI_HIP_BLOCKS_PER_CU")) c->blocks_per_cu = std::max(1, std::min(64, std::atoi(v)));
  if (const char *v = std::getenv("EBI_HIP_BATCH_MIXED")) c->batch_mixed = std::atoi(v) != 0;
  *out = c;
  return EBI_OK;
}

extern "C" int ebi_hip_ctx_destroy(ebi_hip_ctx *ctx) {
  if (!ctx) return EBI_INVALID_ARGUMENT;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  ctx->plan_dev.release();
  if (ctx->plan_pin) hipHostFree(ctx->plan_pin);
  ctx->batch_jobs.release();
  ctx->batch_args.release();
  ctx->gemm_a.release();
  ctx->gemm_b.release();
  ctx->gemm_c.release();
  if (ctx->h_batch) hipHostFree(ctx->h_batch);
  if (ctx->d_flags) hipFree(ctx->d_flags);
  if (ctx->h_flags) hipHostFree(ctx->h_flags);
  if (ctx->h_view) hipHostFree(ctx->h_view);
  for (int i = 0; i < 2; ++i) {
    if (ctx->stage[i]) hipHostFree(ctx->stage[i]);
    if (ctx->stage_ev[i]) hipEventDestroy(ctx->stage_ev[i]);
  }
  if (ctx->ev0) hipEventDestroy(ctx->ev0);
  if (ctx->ev1) hipEventDestroy(ctx->ev1);
  if (ctx->ev_mid) hipEventDestroy(ctx->ev_mid);
  if (ctx->ev_seg) hipEventDestroy(ctx->ev_seg);
  if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
  if (ctx->side) hipStreamDestroy(ctx->side);
  for (hipEvent_t e : ctx->ev_ph)
    if (e) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return EBI_OK;
}

extern "C" const char *ebi_hip_last_error(const ebi_hip_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

extern "C" double ebi_hip_last_device_ms(const ebi_hip_ctx *ctx) { return ctx ? ctx->last_ms : 0.0; }

extern "C" double ebi_hip_last_phase_ms(const ebi_hip_ctx *ctx, int phase) {
  return ctx && phase >= 0 && phase < 4 ? ctx->phase_ms[phase] : 0.0;
}

// Total and per-kernel device time of a decode whose kernels were bracketed
// by ev0, ev_seg, ev_mid, ev1 (phase 0 k_decode, 2 k_decode_seg, 1 the
// Chimp128 / serial kernel).
static void record_decode_times(ebi_hip_ctx *ctx) {
  float a = 0, b = 0, c = 0;
  hipEventElapsedTime(&a, ctx->ev0, ctx->ev_seg);
  hipEventElapsedTime(&c, ctx->ev_seg, ctx->ev_mid);
  hipEventElapsedTime(&b, ctx->ev_mid, ctx->ev1);
  ctx->phase_ms[0] = a;
  ctx->phase_ms[1] = b;
  ctx->phase_ms[2] = c;
  ctx->phase_ms[3] = 0.0;
  ctx->last_ms = (double)a + (double)b + (double)c;
}

extern "C" int ebi_hip_column_load(ebi_hip_ctx *ctx, const uint8_t *file, size_t len, ebi_hip_col **out) {
  EBI_GUARD(ctx);
  if (!file || !out) return EBI_INVALID_ARGUMENT;
  *out = nullptr;
  hipSetDevice(ctx->device);
  ebi_hip_col *c = new (std::nothrow) ebi_hip_col();
  if (!c) return EBI_OOM;
  int rc = load_column(ctx, file, len, c->c);
  if (rc) {
    hipStreamSynchronize(ctx->stream);
    free_column(c->c);
    delete c;
    return rc;
  }
  *out = c;
  return EBI_OK;
}

extern "C" int ebi_hip_column_free(ebi_hip_ctx *ctx, ebi_hip_col *col) {
  EBI_GUARD(ctx);
  if (!col) return EBI_INVALID_ARGUMENT;
  if (col->c.users.load())
    return fail(ctx, EBI_PRECONDITION, "column in use by a queued or running served query or a prepared batch");
  hipStreamSynchronize(ctx->stream);
  free_column(col->c);
  delete col;
  return EBI_OK;
}

extern "C" int ebi_hip_column_info(const ebi_hip_col *col, ebi_file_info *info) {
  if (!col || !info) return EBI_INVALID_ARGUMENT;
  const FileMeta &m = col->c.meta;
  info->number_of_records = m.number_of_records;
  info->number_of_chunks = m.chunks.size();
  info->footer_offset = m.footer_offset;
  info->chunk_option_value = m.chunk_value;
  info->config_capacity = m.capacity;
  info->config_scale = m.scale;
  info->scheme = m.scheme;
  info->chunk_option_kind = m.chunk_kind;
  for (int i = 0; i < 3; ++i) info->version[i] = m.version[i];
  info->reserved = 0;
  return EBI_OK;
}

extern "C" int ebi_hip_column_set_precision(ebi_hip_col *col, int32_t precision) {
  if (!col || precision < -1 || precision > 12) return EBI_INVALID_ARGUMENT;
  col->c.precision = precision;
  return EBI_OK;
}

extern "C" uint64_t ebi_hip_column_device_bytes(const ebi_hip_col *col) { return col ? col->c.device_bytes : 0; }

extern "C" uint64_t ebi_hip_column_compressed_bytes(const ebi_hip_col *col) {
  return col ? col->c.compressed_bytes : 0;
}

extern "C" int ebi_hip_decode_device(ebi_hip_ctx *ctx, const ebi_hip_col *colc, double *d_out) {
  EBI_GUARD(ctx);
  if (!colc || (!d_out && colc->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  int rc = check_chunk_errors(ctx, c, -1);
  if (rc) return rc;
  if ((rc = upload_view(ctx, c, d_out))) return rc;
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  hipEventRecord(ctx->ev0, ctx->stream);
  if ((rc = enqueue_decode(ctx, c))) return rc;
  hipEventRecord(ctx->ev1, ctx->stream);
  if ((rc = sync_and_check(ctx, "decode"))) return rc;
  record_decode_times(ctx);
  return EBI_OK;
}

extern "C" int ebi_hip_decode_per_chunk(ebi_hip_ctx *ctx, const ebi_hip_col *colc, double *d_out) {
  EBI_GUARD(ctx);
  if (!colc || (!d_out && colc->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
  Column &c = const_cast<ebi_hip_col *>(colc)->c;
  int rc = check_chunk_errors(ctx, c, -1);
  if (rc) return rc;
  if ((rc = upload_view(ctx, c, d_out))) return rc;
  if (ctx->batch_args.ensure(4096)) return fail(ctx, EBI_OOM, "batch arguments");
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  hipEventRecord(ctx->ev0, ctx->stream);
  // The plain form: one launch per chunk and a host synchronisation after
  // each, of the same kernels as the batched form over the chunk's own jobs
  // (its segment jobs for the XOR codecs, its tile jobs otherwise; both are
  // contiguous slices of the resident lists, and a segment job's Job::col
  // is absolute, so a slice decodes as in the whole list).
  size_t sj = 0;  // the chunk's first segment job
  for (size_t k = 0; k < c.chunks.size(); ++k) {
    if (c.xor_codec || c.chimp128) {
      size_t sj1 = sj;
      while (sj1 < c.h_seg_jobs.size() && c.h_seg_jobs[sj1].chunk == k) ++sj1;
      ColView v = view_of(c, d_out);
      v.seg_jobs = c.d_seg_jobs + sj;
      v.n_seg_jobs = (uint32_t)(sj1 - sj);
      v.n_tile_jobs = 0;
      uint32_t *pre = reinterpret_cast<uint32_t *>(ctx->h_batch);
      pre[0] = 0;
      pre[1] = v.n_seg_jobs;
      std::memcpy(ctx->h_batch + 256, &v, sizeof v);
      hipMemcpyAsync(ctx->batch_args.p, ctx->h_batch, 256 + sizeof v, hipMemcpyHostToDevice, ctx->stream);
      launch::decode(ctx->stream, (const ColView *)((uint8_t *)ctx->batch_args.p + 256), 1, nullptr, 0,
                     (const uint32_t *)ctx->batch_args.p, v.n_seg_jobs, ctx->d_flags,
                     (uint32_t)ctx->cu_count * 8);
      sj = sj1;
    } else {
      // the chunk's tile jobs are a contiguous slice of the resident list
      const uint32_t a = c.tile_job_first[k], b = c.tile_job_first[k + 1];
      ColView v = view_of(c, d_out);
      v.tile_jobs = c.d_tile_jobs + a;
      v.n_tile_jobs = b - a;
      v.n_seg_jobs = 0;
      uint32_t *pre = reinterpret_cast<uint32_t *>(ctx->h_batch);
      pre[0] = 0;
      pre[1] = b - a;
      std::memcpy(ctx->h_batch + 256, &v, sizeof v);
      hipMemcpyAsync(ctx->batch_args.p, ctx->h_batch, 256 + sizeof v, hipMemcpyHostToDevice, ctx->stream);
      launch::decode(ctx->stream, (const ColView *)((uint8_t *)ctx->batch_args.p + 256), 1,
                     (const uint32_t *)ctx->batch_args.p, b - a, nullptr, 0, ctx->d_flags,
                     (uint32_t)ctx->cu_count * 8, nullptr, false, 1u << c.meta.scheme);
    }
    hipError_t e = stream_wait(ctx->stream);  // the same wait as the batched forms
    if (e != hipSuccess) return hip_fail(ctx, e, "per-chunk decode");
  }
  hipEventRecord(ctx->ev1, ctx->stream);
  if ((rc = sync_and_check(ctx, "per-chunk decode"))) return rc;
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  ctx->last_ms = ms;
  return EBI_OK;
}

extern "C" int ebi_hip_device_alloc(ebi_hip_ctx *ctx, size_t bytes, void **d_ptr) {
  EBI_GUARD(ctx);
  if (!d_ptr) return EBI_INVALID_ARGUMENT;
  hipError_t e = hipMalloc(d_ptr, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
  return EBI_OK;
}

extern "C" int ebi_hip_device_free(ebi_hip_ctx *ctx, void *d_ptr) {
  EBI_GUARD(ctx);
  hipError_t e = hipFree(d_ptr);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
  return EBI_OK;
}

extern "C" int ebi_hip_device_zero(ebi_hip_ctx *ctx, void *d_ptr, size_t bytes) {
  EBI_GUARD(ctx);
  if ((!d_ptr && bytes) || (bytes & 3)) return EBI_INVALID_ARGUMENT;
  const size_t step = (size_t)1 << 33;  // 2^31 words per launch
  for (size_t off = 0; off < bytes; off += step) {
    const size_t n = std::min(bytes - off, step) / 4;
    launch::zero_u32(ctx->stream, reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(d_ptr) + off), (uint32_t)n);
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "device zero");
  return EBI_OK;
}

extern "C" int ebi_hip_copy_to_host(ebi_hip_ctx *ctx, void *h_dst, const void *d_src, size_t bytes) {
  EBI_GUARD(ctx);
  if (!bytes) return EBI_OK;
  hipError_t e = hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "copy to host");
  return EBI_OK;
}

extern "C" int ebi_hip_decode_batch(ebi_hip_ctx *ctx, const ebi_hip_col *const *cols, size_t n_cols,
                                    double *const *d_outs) {
  EBI_GUARD(ctx);
  if (!cols || !d_outs || n_cols == 0 || n_cols > (1u << 16)) return EBI_INVALID_ARGUMENT;
  int rc;
  for (size_t i = 0; i < n_cols; ++i) {
    if (!cols[i] || (!d_outs[i] && cols[i]->c.meta.number_of_records)) return EBI_INVALID_ARGUMENT;
    if ((rc = check_chunk_errors(ctx, cols[i]->c, -1))) return rc;
  }
  // arguments: prefix over tile jobs, prefix over segment jobs, then the
  // ColView array (256-byte aligned)
  const size_t pre_bytes = ((2 * (n_cols + 1) * 4 + 255) / 256) * 256;
  const size_t bytes = pre_bytes + n_cols * sizeof(ColView);
  if (bytes > ctx->h_batch_cap) {
    hipHostFree(ctx->h_batch);
    ctx->h_batch = nullptr;
    ctx->h_batch_cap = 0;
    if (hipHostMalloc((void **)&ctx->h_batch, bytes, hipHostMallocDefault) != hipSuccess)
      return fail(ctx, EBI_OOM, "batch staging");
    ctx->h_batch_cap = bytes;
  }
  if (ctx->batch_args.ensure(bytes)) return fail(ctx, EBI_OOM, "batch arguments");
  uint32_t *pre = reinterpret_cast<uint32_t *>(ctx->h_batch);
  uint32_t *preseg = pre + (n_cols + 1);
  ColView *views = reinterpret_cast<ColView *>(ctx->h_batch + pre_bytes);
  uint64_t jobs = 0, segs = 0;
  uint32_t codecs = 0;
  for (size_t i = 0; i < n_cols; ++i) {
    const Column &c = cols[i]->c;
    pre[i] = (uint32_t)jobs;
    preseg[i] = (uint32_t)segs;
    views[i] = view_of(c, d_outs[i]);
    jobs += c.n_tile_jobs;
    segs += c.n_seg_jobs;
    codecs |= 1u << c.meta.scheme;
  }
  pre[n_cols] = (uint32_t)jobs;
  preseg[n_cols] = (uint32_t)segs;
  if (jobs >= (1ull << 32) || segs >= (1ull << 32))
    return fail(ctx, EBI_UNSUPPORTED, "batch too large");
  uint8_t *d = (uint8_t *)ctx->batch_args.p;
  hipError_t e = hipMemcpyAsync(d, ctx->h_batch, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "batch upload");
  launch::zero_u32(ctx->stream, ctx->d_flags, 2);
  hipEventRecord(ctx->ev0, ctx->stream);
  const ColView *dv = (const ColView *)(d + pre_bytes);
  launch::decode(ctx->stream, dv, (uint32_t)n_cols, (const uint32_t *)d, (uint32_t)jobs,
                 (const uint32_t *)d + (n_cols + 1), (uint32_t)segs, ctx->d_flags,
                 (uint32_t)ctx->cu_count * ctx->blocks_per_cu, ctx->ev_seg, false, codecs);
  hipEventRecord(ctx->ev_mid, ctx->stream);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(ctx, e, "batch launch");
  hipEventRecord(ctx->ev1, ctx->stream);
  if ((rc = sync_and_check(ctx, "batch decode"))) return rc;
  float a = 0, b = 0, c = 0, t = 0;
  hipEventElapsedTime(&a, ctx->ev0, ctx->ev_seg);
  hipEventElapsedTime(&c, ctx->ev_seg, ctx->ev_mid);
  hipEventElapsedTime(&b, ctx->ev_mid, ctx->ev1);
  hipEventElapsedTime(&t, ctx->ev0, ctx->ev1);
  ctx->phase_ms[0] = a;
  ctx->phase_ms[1] = b;
  ctx->phase_ms[2] = c;
  ctx->phase_ms[3] = 0.0;
  ctx->last_ms = t;
  return EBI_OK;
}
