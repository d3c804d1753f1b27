// Query plans: one Decoder query (materialize, filter, filter_materialize,
// sum, min, max; ebi/src/api/decoder.rs:202-439) over a resident column,
// split into an asynchronous device part and a host finish, so that the
// synchronous C entry points and the serving scheduler run the same code.
//
// Plans are issued in groups (the serving scheduler's batches; a
// synchronous call is a group of one).  A group costs a fixed number of
// kernel launches whatever its size: every launch covers all of its plans
// through a device array of plan descriptors (PlanDesc) and prefix arrays
// mapping flat work items to plans.
//
// Life of a plan:
//   plan_init   validate the request, derive what it runs (host only)
//   plan_need   device and pinned bytes of its scratch
//   plan_bind   carve that scratch out of caller-provided memory
//   plan_stage  host staging into pinned memory (input bitmask copy)
//   group_enqueue (per group, on one stream): one upload of the group block
//               (views, descriptors, prefixes, fills), each plan's bitmask /
//               predicate upload, then the group kernels: scratch fills,
//               chunk selection, ONE batched decode launch over every plan
//               that needs decoded records, the record-tile pass (filters
//               and aggregates in one pass; BUFF / DeltaSprintz on their
//               codes), the fused decode + query pass of the XOR codecs,
//               the per-chunk and final folds (result on the device) and
//               the stream compaction of materializing plans; then each
//               FILTER plan's bitset and one copy of every plan's result
//               record back
//   plan_enqueue_values  after the device work: copy the materialized
//               records (count known only then) into the caller's buffer
//   plan_finish results into the caller's buffers
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/ebi_hip.h"
#include "column.hpp"

namespace ebi {

enum PlanOp : int {
  kOpMaterialize = EBI_OP_MATERIALIZE,
  kOpFilter = EBI_OP_FILTER,
  kOpFilterMaterialize = EBI_OP_FILTER_MATERIALIZE,
  kOpSum = EBI_OP_SUM,
  kOpMin = EBI_OP_MIN,
  kOpMax = EBI_OP_MAX,
  kOpFilterSum = EBI_OP_FILTER_SUM,
  kOpDecode = EBI_OP_DECODE,
};

struct PlanNeed {
  uint64_t dev = 0;
  uint64_t pin = 0;
};

struct Plan {
  ebi_query q{};
  Column *c = nullptr;
  bool insitu = false;     // BUFF / DeltaSprintz operate on codes
  bool decode = false;     // needs the column's records decoded into `values`
  bool has_mask = false;
  bool direct = false;     // the bitmask and the filter result move between the caller's host
                           // buffers and the device directly (no staging of them)
  bool compact = false;    // materialize through the bitset and compaction (a bitmask, or a
                           // column whose chunks end at NaNs: its records are not contiguous)
  bool cpred = false;      // per-chunk predicate parameters (computed by the selection pass)
  uint32_t flags = 0;      // PlanFlags (device.hpp)
  int filter_mode = 0;     // record-tile filter mode (0 select, 1 code domain / f64, 2 f64)
  int tile_kind = 0;       // SpanKind (device.hpp): which record-tile kernel runs the plan
  uint64_t words = 0;      // ceil(N / 64)
  uint32_t nc = 0;         // chunks
  uint32_t n_blocks = 0;   // compaction blocks
  // device scratch
  uint64_t *mask = nullptr, *bits = nullptr, *acc = nullptr, *block = nullptr;
  uint32_t *sel = nullptr;
  ChunkPred *d_cpred = nullptr;
  double *partials = nullptr, *chunk_sum = nullptr, *values = nullptr, *out = nullptr;
  // pinned staging
  uint64_t *h_mask = nullptr, *h_bits = nullptr;
  // in the group block (set by group_bind)
  const ColView *d_view = nullptr;
  PlanResult *d_result = nullptr, *h_result = nullptr;
  uint64_t n_out = 0;      // records materialized
  // result
  int status = EBI_OK;
  std::string error;
};

// Validates `q` (0 or an ebi_status; sets p.error).  Host only.
int plan_init(Plan &p, const ebi_query &q);
PlanNeed plan_need(const Plan &p);
void plan_bind(Plan &p, uint8_t *dev, uint8_t *pin);
void plan_stage(Plan &p);
// Enqueues the D2H copy of materialized records once the device work has
// completed (reads the group's result record).  Returns false if nothing
// was enqueued.
bool plan_enqueue_values(Plan &p, hipStream_t s, hipError_t *err);
// Results into the caller's buffers; returns the plan's status.
int plan_finish(Plan &p);

// A set of plans issued together on one stream: their views, descriptors,
// prefixes, fills and result records live in one block (device + pinned
// mirror) of group_block_bytes(n) bytes.
struct PlanGroup {
  std::vector<Plan *> plans;
  uint8_t *dev = nullptr;
  uint8_t *pin = nullptr;
  size_t cap = 0;               // plans the block was laid out for
  uint32_t *d_flags = nullptr;  // [0] decode error flags
  uint32_t *h_flags = nullptr;
};
uint64_t group_block_bytes(size_t n_plans);
// Binds the block (laid out for `cap` plans, cap >= plans.size()).
void group_bind(PlanGroup &g, uint8_t *dev, uint8_t *pin, size_t cap);
// Issues the whole group; the persistent decode grids are cu_count *
// blocks_per_cu blocks.
hipError_t group_enqueue(PlanGroup &g, hipStream_t s, int cu_count, uint32_t blocks_per_cu = 8);

}  // namespace ebi
