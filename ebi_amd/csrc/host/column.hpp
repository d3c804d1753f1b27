// Host-side state of libebi_hip: the device context and resident columns.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../device.hpp"
#include "format.hpp"

struct ebi_hip_ctx;

namespace ebi {

// Grow-only device buffer owned by a context.
struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);  // 0 or hip error
  void release();
};

// Per-chunk host copy of the parsed codec header (also uploaded in DevChunk).
struct HostChunk {
  ChunkGeom geom;
  DevChunk dev;       // with absolute device addresses filled in
  uint32_t err = 0;   // DevError flags found at load
  uint32_t n_tiles = 0;
  uint32_t n_segs = 0;
};

struct Column {
  FileMeta meta;
  std::vector<HostChunk> chunks;
  uint64_t compressed_bytes = 0;  // sum of chunk sizes in the file
  uint64_t device_bytes = 0;      // bytes of the device body allocation
  // device allocations
  uint8_t *d_body = nullptr;
  DevChunk *d_chunks = nullptr;
  int64_t *d_tile_index = nullptr;
  XorCkpt *d_ckpt = nullptr;
  uint4 *d_ilv = nullptr;    // lane-interleaved segment streams (Gorilla, Chimp, Elf)
  uint64_t *d_c128_ring = nullptr;  // Chimp128: per segment job the 128 values before it
  uint64_t ilv_rows = 0;     // 1 KiB rows of d_ilv
  uint32_t *d_chunk_err = nullptr;
  uint32_t *d_chunk_end = nullptr;  // load-time passes: records before a chunk's first NaN
  ColView *d_view = nullptr;  // slot used by single-column launches
  uint32_t *d_prefix = nullptr;  // {0, tile jobs, 0, serial jobs, 0, chunks, 0, seg jobs}: single-column launches
  // resident job lists
  Job *d_tile_jobs = nullptr;
  uint32_t n_tile_jobs = 0;
  Job *d_serial_jobs = nullptr;  // Gorilla/Chimp/Elf/Chimp128, lane per chunk
  uint32_t n_serial_jobs = 0;
  Job *d_single_jobs = nullptr;  // one single-lane serial job per XOR chunk (comparator)
  Job *d_seg_jobs = nullptr;     // Gorilla/Chimp/Elf checkpoint segments
  Job *d_rec_jobs = nullptr;     // record tiles of every chunk (select filter)
  uint32_t n_rec_jobs = 0;
  Job *d_span_jobs = nullptr;    // up to kSpanTiles record tiles of one chunk (query plans)
  uint32_t n_span_jobs = 0;
  Job *d_field_spans = nullptr;  // RLE: up to kSpanTiles field tiles of one chunk (per-run queries)
  uint32_t span_tiles = kSpanTiles, field_span_tiles = kSpanTiles;  // chosen span lengths
  uint32_t n_field_spans = 0;
  bool rle_direct = false;       // RLE whose chunks have no more fields than records: queried per run
  uint64_t n_rtiles = 0;
  uint32_t n_seg_jobs = 0;
  std::vector<Job> h_tile_jobs, h_serial_jobs, h_seg_jobs;
  std::vector<uint32_t> tile_job_first;  // per chunk: first tile job index (comparator)
  std::vector<uint16_t> rle_steps;       // RLE: output steps (kTile records) of each field tile
  uint64_t tile_index_entries = 0, ckpt_entries = 0;
  bool xor_codec = false;
  bool chimp128 = false;
  // some chunk ends at a NaN before its footer record count (its DevChunk n
  // is shortened to the records before the NaN): the column's records are no
  // longer contiguous, so a whole-column materialize compacts them
  bool truncated = false;
  // Decoder::with_precision (api/decoder.rs:159-161): -1 full, else 0..12;
  // BUFF decodes read only the planes that precision needs
  int32_t precision = -1;
  // plans and prepared batches that use the column (column_free refuses
  // while this is nonzero)
  std::atomic<uint32_t> users{0};
};

}  // namespace ebi

struct ebi_hip_col {
  ebi::Column c;
};

struct ebi_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_seg = nullptr;  // between the tile kernel and the segment kernel of a decode
  hipEvent_t ev_mid = nullptr;  // between the segment kernel and the Chimp128 kernel of a decode
  // captured batches: the Uncompressed copy runs on `side` beside the segment kernel
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_ph[4] = {nullptr, nullptr, nullptr, nullptr};  // phase boundaries of matmul calls
  double last_ms = 0.0;
  double phase_ms[4] = {0.0, 0.0, 0.0, 0.0};  // of the last decode_device / decode_batch / matmul call
  std::string last_error;
  std::recursive_mutex mu;
  uint32_t *d_flags = nullptr;  // [0] error flags, [1] job counter, [2..] scratch counters
  uint32_t *h_flags = nullptr;  // pinned mirror
  ebi::ColView *h_view = nullptr;  // pinned staging for ColView uploads
  size_t h_view_cap = 0;
  ebi::DevBuf batch_args;       // device ColView array + prefixes for batches
  uint8_t *h_batch = nullptr;     // pinned staging for batch arguments
  size_t h_batch_cap = 0;
  ebi::DevBuf plan_dev;         // device scratch of the synchronous query calls
  uint8_t *plan_pin = nullptr;  // pinned staging of the synchronous query calls
  size_t plan_pin_cap = 0;
  ebi::DevBuf batch_jobs;
  ebi::DevBuf gemm_a, gemm_b, gemm_c;  // decoded operands and product of the matmul entry points
  int cu_count = 256;
  // launch shape of the decode kernels (defaults; a prepared batch tunes its own)
  uint32_t blocks_per_cu = 8;   // persistent k_decode grid = cu_count * blocks_per_cu
  uint32_t batch_mixed = 0;     // untuned prepared batches: 1 = one k_decode_mix launch
  std::map<std::string, ebi_batch_config> tune_cache;  // prepared-batch shapes by batch signature
  int xor_mode = 0;  // 0: checkpoint segments, 1: lane-serial chunks (Gorilla, Chimp, Elf)
  int xindex_serial = 0;  // 1: XOR load index by one lane per chunk instead of the parallel index
  // pinned double buffer of the column upload (allocated at the first load)
  uint8_t *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  uint8_t *noalloc_dev = nullptr;  // this device's scratch of ebi_hip_query_noalloc
};
