/*
 * ebi_hip.h -- C ABI of the MI355X (gfx950) chunked-decode and in-situ query
 * library for EBI1 files (libebi_hip.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, throws no
 * C++ exception across the boundary and returns an ebi_status.  Host buffers
 * are owned by the caller and borrowed for the duration of one call (the
 * shape of cuda-binding/src/lib.rs:59-140 in the reference); device memory
 * is owned by the library through the ctx / col handles.  A ctx serialises
 * its calls on one HIP stream, so calls from several threads on one ctx are
 * safe but do not overlap; the serving entry points (ebi_hip_serve_*) are the
 * concurrent path.
 *
 * Bitsets are dense little-endian u64 words over GLOBAL record ids: record r
 * is bit (r % 64) of word (r / 64).  They replace the reference's
 * RoaringBitmap of global ids (ebi/src/decoder/query.rs:3); the reference's
 * own BUFF filter builds its result from lsb0 bytes the same way
 * (decoder/chunk_reader/buff/filter_cmp.rs:532-541).
 *
 * Reference citations are relative to the lemolatoon/ebi repository.
 */
#ifndef EBI_HIP_H
#define EBI_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  They map one-to-one onto decoder::DecoderError
 * (ebi/src/decoder/error.rs:6-21) plus the device-side conditions. */
typedef enum ebi_status {
    EBI_OK = 0,
    EBI_UNEXPECTED_END = 1, /* DecoderError::UnexpectedEndOfChunk */
    EBI_CONVERSION = 2,     /* DecoderError::ConversionError (bad scheme / field byte) */
    EBI_PRECONDITION = 3,   /* DecoderError::PreconditionsNotMet */
    EBI_HIP = 4,            /* a HIP runtime call failed (was CudaError) */
    EBI_OOM = 5,            /* device or host allocation failed */
    EBI_REJECTED = 6,       /* serving admission refused the request */
    EBI_INVALID_ARGUMENT = 7,
    EBI_UNSUPPORTED = 8,    /* scheme outside the accelerated set */
    EBI_PENDING = 9         /* serving: the ticket has not completed yet */
} ebi_status;

/* Compression scheme ids, format.rs:185-199. */
enum {
    EBI_SCHEME_UNCOMPRESSED = 0,
    EBI_SCHEME_RLE = 1,
    EBI_SCHEME_BUFF = 2,
    EBI_SCHEME_GORILLA = 3,
    EBI_SCHEME_CHIMP = 4,
    EBI_SCHEME_CHIMP128 = 5,
    EBI_SCHEME_ELF_ON_CHIMP = 6,
    EBI_SCHEME_ELF = 7,
    EBI_SCHEME_DELTA_SPRINTZ = 8
};

/* ChunkOption kinds, format.rs:173-178 / encoder.rs:26-64. */
enum { EBI_CHUNK_FULL = 0, EBI_CHUNK_RECORD_COUNT = 1, EBI_CHUNK_BYTE_SIZE = 2 };

/* Predicate, decoder/query.rs:347-446.  kind: 0 Eq(value), 1 Ne(value),
 * 2 Range(start, end); bound kinds: 0 None, 1 Inclusive, 2 Exclusive. */
enum { EBI_PRED_EQ = 0, EBI_PRED_NE = 1, EBI_PRED_RANGE = 2 };
enum { EBI_BOUND_NONE = 0, EBI_BOUND_INCLUSIVE = 1, EBI_BOUND_EXCLUSIVE = 2 };
typedef struct ebi_pred {
    int32_t kind;
    int32_t start_kind;
    int32_t end_kind;
    int32_t reserved;
    double value; /* Eq / Ne operand */
    double start; /* Range start operand */
    double end;   /* Range end operand */
} ebi_pred;

/* ------------------------------------------------------------------ */
/* Write path (CPU): the reference's Encoder::encode                   */
/* (ebi/src/api/encoder.rs:140-178, encoder.rs:169-405).               */
/* ------------------------------------------------------------------ */
typedef struct ebi_encoder_config {
    int32_t scheme;       /* EBI_SCHEME_* (0..8) */
    int32_t chunk_kind;   /* EBI_CHUNK_* */
    uint64_t chunk_value; /* records (RECORD_COUNT) or bytes (BYTE_SIZE) */
    uint64_t capacity;    /* Capacity config field; 0 -> 8192 (compressor.rs:465) */
    uint64_t scale;       /* BUFF / DeltaSprintz scale; ignored otherwise */
    int32_t n_threads;    /* chunks are encoded on this many host threads; 0 -> 1 */
    int32_t reserved;
} ebi_encoder_config;

/* Encode n f64 values into one EBI1 file.  *out is allocated by the library
 * (free it with ebi_buffer_free).  Bytes are identical for any n_threads. */
int ebi_encode(const double *values, size_t n, const ebi_encoder_config *cfg,
               uint8_t **out, size_t *out_len);
void ebi_buffer_free(void *p);

/* Seeded synthetic column (benchmark and test workloads, not a reference
 * interface).  kind 0: random walk (p0 = x0, p1 = sigma); 1: uniform
 * [p0, p1); 2: step series with change probability p0; 3: 100 sin(2 pi i /
 * p0) + p1 N(0,1).  Values are rounded half away from zero to `decimals`. */
int ebi_synth_column(int kind, uint64_t n, uint64_t seed, int decimals, double p0, double p1,
                     double p2, int threads, double *out);

/* ------------------------------------------------------------------ */
/* File metadata (CPU): FileReader::fetch_header / fetch_footer         */
/* (ebi/src/decoder.rs:131-220) and GeneralChunkHandle geometry         */
/* (decoder.rs:283-352).                                                */
/* ------------------------------------------------------------------ */
typedef struct ebi_file_info {
    uint64_t number_of_records;
    uint64_t number_of_chunks;
    uint64_t footer_offset;
    uint64_t chunk_option_value;
    uint64_t config_capacity;
    uint64_t config_scale;
    int32_t scheme;
    int32_t chunk_option_kind;
    uint16_t version[3];
    uint16_t reserved;
} ebi_file_info;

int ebi_file_parse(const uint8_t *file, size_t len, ebi_file_info *info);
/* Per-chunk geometry; arrays hold info.number_of_chunks entries (nullable). */
int ebi_file_chunks(const uint8_t *file, size_t len, uint64_t *physical_offset,
                    uint64_t *chunk_size, uint64_t *logical_offset,
                    uint64_t *number_of_records);

/* ------------------------------------------------------------------ */
/* Device context and resident columns                                 */
/* ------------------------------------------------------------------ */
typedef struct ebi_hip_ctx ebi_hip_ctx;
typedef struct ebi_hip_col ebi_hip_col;

/* Replaces cuda_binding::Context::new_at (cuda-binding/src/lib.rs:41-53). */
int ebi_hip_ctx_create(int device, ebi_hip_ctx **out);
int ebi_hip_ctx_destroy(ebi_hip_ctx *ctx);
/* Message of the last failing call on this ctx ("" if none). */
const char *ebi_hip_last_error(const ebi_hip_ctx *ctx);

/* Parse header + footer (Decoder::new, api/decoder.rs:136-156), build the
 * chunk-descriptor table and upload chunk bytes; the column stays resident
 * until ebi_hip_column_free. */
int ebi_hip_column_load(ebi_hip_ctx *ctx, const uint8_t *file, size_t len,
                        ebi_hip_col **out);
int ebi_hip_column_free(ebi_hip_ctx *ctx, ebi_hip_col *col);
int ebi_hip_column_info(const ebi_hip_col *col, ebi_file_info *info);
/* Decoder::with_precision (api/decoder.rs:159-161): decodes of a BUFF column
 * (materialize, filter_materialize, the device decodes and the matmul
 * operands) use decode_with_precision (chunk_reader/buff/
 * controlled_precision.rs:5-75) with this many decimal digits, reading only
 * the byte planes those digits need; in-situ filters and aggregates are
 * unaffected, as in the reference.  precision -1 restores full precision;
 * other codecs ignore it.  0..12 (PRECISION_MAP's range), else
 * EBI_INVALID_ARGUMENT.  Prepared batches keep the precision the column had
 * when they were created. */
int ebi_hip_column_set_precision(ebi_hip_col *col, int32_t precision);
/* Bytes of chunk data resident for the column (excluding padding). */
uint64_t ebi_hip_column_compressed_bytes(const ebi_hip_col *col);
/* Device bytes the resident column holds: chunk bodies, descriptors, job
 * lists and the load-time index (XOR checkpoints, the lane-interleaved
 * stream copy, Chimp128 ring checkpoints, tile prefixes).  No counterpart in
 * the reference (its Decoder reads the file in host memory); for capacity
 * planning.  ebi_hip_last_phase_ms after ebi_hip_column_load: phase 0 =
 * parse, repack and upload, 1 = the index pass, 2 = the interleaved copy /
 * ring checkpoints (host clock, each phase ends in a stream sync). */
uint64_t ebi_hip_column_device_bytes(const ebi_hip_col *col);

/* Query entry points.  They mirror Decoder::{materialize, filter,
 * filter_materialize, sum, min, max} (api/decoder.rs:202-439):
 *   bitset   -- nullable input bitmask over global ids, ceil(N/64) words;
 *   chunk_id -- -1 for every chunk, else ChunkId (api/decoder.rs:97-106, chunk_reader :726-737);
 * chunks whose record range has no bit set are pruned exactly as the
 * reference does (api/decoder.rs:222-226). */
int ebi_hip_materialize(ebi_hip_ctx *ctx, const ebi_hip_col *col,
                        const uint64_t *bitset, int64_t chunk_id,
                        double *out, size_t *n_out);
int ebi_hip_filter(ebi_hip_ctx *ctx, const ebi_hip_col *col, const ebi_pred *pred,
                   const uint64_t *bitset, int64_t chunk_id, uint64_t *out_bitset);
int ebi_hip_filter_materialize(ebi_hip_ctx *ctx, const ebi_hip_col *col,
                               const ebi_pred *pred, const uint64_t *bitset,
                               int64_t chunk_id, double *out, size_t *n_out);
int ebi_hip_sum(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset,
                int64_t chunk_id, double *out);
int ebi_hip_min(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset,
                int64_t chunk_id, double *out);
int ebi_hip_max(ebi_hip_ctx *ctx, const ebi_hip_col *col, const uint64_t *bitset,
                int64_t chunk_id, double *out);

/* One query as data: the same six operations, for ebi_hip_query and the
 * serving loop.  Output pointers are the caller's host buffers. */
enum {
    EBI_OP_MATERIALIZE = 0,        /* out_values (room for N), n_out */
    EBI_OP_FILTER = 1,             /* out_bitset (ceil(N/64) words) */
    EBI_OP_FILTER_MATERIALIZE = 2, /* out_values (room for N), n_out */
    EBI_OP_SUM = 3,                /* out_scalar */
    EBI_OP_MIN = 4,
    EBI_OP_MAX = 5,
    /* sum over the records the predicate selects: the same result as FILTER
     * followed by SUM with that bitset (bit for bit), without the round trip */
    EBI_OP_FILTER_SUM = 6,         /* out_scalar */
    /* whole-column decode into the DEVICE buffer out_values (N doubles),
     * as ebi_hip_decode_device */
    EBI_OP_DECODE = 7
};
typedef struct ebi_query {
    int32_t op;              /* EBI_OP_* */
    int32_t reserved;
    const ebi_hip_col *col;
    ebi_pred pred;           /* FILTER, FILTER_MATERIALIZE, FILTER_SUM */
    const uint64_t *bitset;  /* nullable input bitmask over global ids */
    int64_t chunk_id;        /* -1 for every chunk */
    double *out_values;
    uint64_t *out_bitset;
    double *out_scalar;
    size_t *n_out;
} ebi_query;

/* Runs one query synchronously on the ctx stream (the ebi_hip_materialize
 * ... ebi_hip_max entry points are this call with the matching op). */
int ebi_hip_query(ebi_hip_ctx *ctx, const ebi_query *q);

/* ebi_hip_query without allocations: no counterpart in the reference; for
 * callers that cannot allocate device or pinned memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * query's device scratch is a fixed array of 64 MiB per GPU, held from the
 * first context on that GPU until the process ends, and its host staging a
 * fixed 4 MiB region per device (devices 0-15); the bitmask and the FILTER
 * result move between the caller's buffers and the device directly.  A query
 * that needs more scratch (about: 2 x ceil(N/64) x 8 bytes of bitsets, 16
 * bytes per 512 records of partial sums, ~150 bytes per chunk, and for
 * MATERIALIZE / FILTER_MATERIALIZE of codecs decoded to HBM 8 or 16 bytes
 * per record) fails with EBI_INVALID_ARGUMENT and a message naming both
 * sizes.  Every call on one GPU shares its scratch (each call clears what it
 * needs), so calls that overlap on one GPU -- from two contexts on the same
 * device -- race; calls on one ctx are serialised as for every entry point.
 * Results equal ebi_hip_query's. */
int ebi_hip_query_noalloc(ebi_hip_ctx *ctx, const ebi_query *q);

/* Device-resident decode: writes all N records of the column to the device
 * buffer d_out (N doubles) on the ctx stream and waits for completion.
 * This is the whole-column decode the batched and DGEMM paths build on. */
int ebi_hip_decode_device(ebi_hip_ctx *ctx, const ebi_hip_col *col, double *d_out);

/* Batched decode of several resident columns in ONE persistent launch over
 * a device-side chunk-descriptor queue (mixed codecs, uneven chunks).
 * d_outs[i] receives column i. */
int ebi_hip_decode_batch(ebi_hip_ctx *ctx, const ebi_hip_col *const *cols,
                         size_t n_cols, double *const *d_outs);

/* Prepared batch: the same decode for a fixed set of columns and outputs,
 * replayed many times.  The job-space arguments are uploaded once; with
 * autotune != 0 the persistent grid size, (when the batch holds both tile
 * and segment jobs) the mixed or two-kernel launch and the tiles per
 * tile-codec job are timed on this batch and the fastest kept, and
 * remembered per batch signature (codec, chunk-size and bit-width bucket of
 * each column) on the context, so an equal batch reuses it without timing
 * (without autotune: the context's defaults, 4 tiles per job, the two
 * kernels unless EBI_HIP_BATCH_MIXED=1 was set when it was created); the
 * launch sequence is captured in a hipGraph.  Columns and outputs must outlive the batch.  No
 * counterpart in the reference. */
typedef struct ebi_hip_batch ebi_hip_batch;
typedef struct ebi_batch_config {
    uint32_t blocks_per_cu; /* persistent decode grids = CUs x blocks_per_cu */
    uint32_t mixed;         /* 1: tile and segment jobs in one launch (k_decode_mix),
                               0: k_decode then k_decode_seg */
    uint32_t tiles_per_job; /* 1, 2, 4 or 8: 512-record tiles (RLE: output steps) per
                               tile-codec decode job */
    uint32_t cached;        /* 1: taken from the context's tuning cache (an equal
                               batch was tuned before), no timing */
    double tuned_ms;        /* device ms of the chosen shape (0 if not tuned) */
} ebi_batch_config;
int ebi_hip_batch_create(ebi_hip_ctx *ctx, const ebi_hip_col *const *cols, size_t n_cols,
                         double *const *d_outs, int autotune, ebi_hip_batch **out);
/* Forces a launch shape (and re-captures the graph): tiles_per_job 1/2/4/8,
 * blocks_per_cu 1..64, mixed 0/1; EBI_INVALID_ARGUMENT otherwise.  No
 * counterpart in the reference (tuning and shape sweeps). */
int ebi_hip_batch_config_set(ebi_hip_ctx *ctx, ebi_hip_batch *b, const ebi_batch_config *cfg);
/* Replays the batch and waits.  eager = 0: one hipGraph launch
 * (ebi_hip_last_device_ms only); eager != 0: the same kernels launched one by
 * one with per-kernel events (ebi_hip_last_phase_ms as for
 * ebi_hip_decode_batch). */
int ebi_hip_batch_run(ebi_hip_ctx *ctx, ebi_hip_batch *b, int eager);
int ebi_hip_batch_config_get(const ebi_hip_batch *b, ebi_batch_config *cfg);
int ebi_hip_batch_destroy(ebi_hip_ctx *ctx, ebi_hip_batch *b);

/* Comparator: the plain per-chunk-launch form -- one launch of the decode
 * kernel per chunk (over that chunk's own tile or segment jobs) and a host
 * synchronisation after each. */
int ebi_hip_decode_per_chunk(ebi_hip_ctx *ctx, const ebi_hip_col *col, double *d_out);

/* Decoder::matmul_cuda (api/decoder.rs:659-724, over
 * cuda_binding::Context::gemm, cuda-binding/src/lib.rs:59-140): the column
 * holds batch = N / (data_rows * data_cols) row-major data matrices A_b
 * (data_rows x data_cols) back to back; `target` is the row-major host matrix
 * B (target_rows x target_cols); `out` receives batch row-major products
 * C_b = A_b * B (data_rows x target_cols each).  EBI_PRECONDITION when
 * data_cols != target_rows, a dimension is 0, or N is not a multiple of the
 * data matrix size.  A data matrix may span chunks (the reference requires
 * each one inside a chunk, chunk_reader.rs:261-280).  Decode and DGEMM run
 * on the device (FP64 matrix cores); phases for ebi_hip_last_phase_ms:
 * 0 decode, 1 target upload, 2 DGEMM. */
int ebi_hip_matmul(ebi_hip_ctx *ctx, const ebi_hip_col *col, const double *target,
                   size_t target_rows, size_t target_cols, size_t data_rows,
                   size_t data_cols, double *out);

/* C = A * B with A (m x k) the whole decoded column col_a and B (k x n) the
 * whole decoded column col_b, both row-major (e.g. matrices stored one row
 * per chunk).  C (m x n, row-major) goes to c_device (a device buffer) if
 * given, else to c_host.  EBI_PRECONDITION unless the columns hold exactly
 * m*k and k*n records.  Phases: 0 decode A, 1 decode B, 2 DGEMM. */
int ebi_hip_gemm_decoded(ebi_hip_ctx *ctx, const ebi_hip_col *col_a,
                         const ebi_hip_col *col_b, size_t m, size_t n, size_t k,
                         double *c_host, double *c_device);

/* The DGEMM alone on device buffers (row-major, C = A * B); phase 0. */
int ebi_hip_dgemm_device(ebi_hip_ctx *ctx, const double *d_a, const double *d_b,
                         double *d_c, size_t m, size_t n, size_t k);

/* ------------------------------------------------------------------ */
/* ML queries on compressed columns (decoded on the device).  Every fold */
/* the reference computes as a sequential f64 sum runs in the same order */
/* on the GPU, so results are bit-identical to it.                       */
/* ------------------------------------------------------------------ */
typedef struct ebi_knn_result {
    uint64_t index;   /* usize::MAX (UINT64_MAX) when no vector was accepted */
    double distance;  /* squared distance; INFINITY when none */
} ebi_knn_result;

/* Decoder::knn1 (api/decoder.rs:469-533): the column read as consecutive
 * vectors of `len` records; the nearest to `target` by squared distance,
 * with the reference's walk reproduced exactly: a vector begun at the end
 * of a chunk is completed from the next chunk's first records (and then
 * compared with `<`), and that chunk's own vectors still start at its record
 * 0; vectors inside a chunk are compared with `<=`.  len 0 (which never ends
 * the reference's loop) is EBI_INVALID_ARGUMENT. */
int ebi_hip_knn1(ebi_hip_ctx *ctx, const ebi_hip_col *col, const double *target, size_t len,
                 ebi_knn_result *out);
/* GeneralChunkReader::{distance_squared, dot_product} (chunk_reader.rs:233-260,
 * decoder/query.rs:157-192) on chunk `chunk_id`: the records from
 * offset_in_chunk (clamped to the chunk's decoded records) zipped with
 * target. */
int ebi_hip_distance_squared(ebi_hip_ctx *ctx, const ebi_hip_col *col, uint64_t chunk_id,
                             size_t offset_in_chunk, const double *target, size_t len, double *out);
int ebi_hip_dot_product(ebi_hip_ctx *ctx, const ebi_hip_col *col, uint64_t chunk_id,
                        size_t offset_in_chunk, const double *target, size_t len, double *out);
/* Decoder::matmul (api/decoder.rs:606-656): out = the batch row-major
 * products of the column's data matrices (data_rows x data_cols, row-major,
 * back to back) with the target matrix given column-first (target_cols
 * columns of target_rows values); each element is the reference's
 * sequential dot product (bit-identical), each batch read where the
 * reference's walk reads it.  EBI_PRECONDITION as the reference
 * (data_cols != target_rows, N not a multiple of the data matrix size, or
 * more matrices than the chunks hold). */
int ebi_hip_matmul_dot(ebi_hip_ctx *ctx, const ebi_hip_col *col, const double *target,
                       size_t target_rows, size_t target_cols, size_t data_rows, size_t data_cols,
                       double *out);

/* ------------------------------------------------------------------ */
/* Multi-column aggregations (Aggregation / Expr, api/decoder.rs:755-917,
 * decoder/expr.rs): expressions over the records of several columns with
 * the same chunking, in postfix form. */
/* ------------------------------------------------------------------ */
enum {
    EBI_EXPR_INDEX = 0,   /* Expr::Index(index): the column cols[index] */
    EBI_EXPR_LITERAL = 1, /* Expr::Literal(literal) */
    EBI_EXPR_ADD = 2,     /* Op::Add of the two entries below (lhs deeper) */
    EBI_EXPR_SUB = 3,
    EBI_EXPR_MUL = 4,
    EBI_EXPR_DIV = 5
};
typedef struct ebi_expr_node {
    int32_t kind;
    uint32_t index;
    double literal;
} ebi_expr_node;
/* AggregationKind (api/decoder.rs:756-761). */
enum { EBI_AGG_NONE = 0, EBI_AGG_MAX = 1, EBI_AGG_MIN = 2, EBI_AGG_SUM = 3 };
typedef struct ebi_aggregation {
    int32_t kind;                /* EBI_AGG_* */
    uint32_t n_nodes;            /* postfix program, at most 32 nodes, stack depth 16 */
    const ebi_expr_node *nodes;
    double *out_scalar;          /* MAX / MIN / SUM */
    double *out_values;          /* NONE: the expression's records, chunk by chunk */
    size_t capacity;             /* NONE: room in out_values */
    size_t *n_out;               /* NONE: records written */
} ebi_aggregation;
/* Aggregation::compute_multiple (api/decoder.rs:874-917) over cols[0..n_cols)
 * with the optional bitmask (global ids, words for the longest column).  An
 * aggregation of one bare column runs that column's own sum / min / max
 * (in-situ for BUFF and DeltaSprintz); every other one is evaluated per
 * chunk i < decoder 0's chunk count, record k of chunk i taking the k-th
 * selected record of chunk i of each column it reads (the reference's
 * zips), up to the shortest.  SUM: within 1e-12 relative of a compensated
 * sum (double-double per 512 records); MAX / MIN: f64::max / f64::min, from
 * -INFINITY / INFINITY.  EBI_PRECONDITION for a literal alone or a binary
 * node of two literals (decoder/expr.rs:73, :97) or a column with fewer
 * chunks than cols[0]; EBI_UNSUPPORTED for NONE over a bare column (the
 * reference's todo!()). */
int ebi_hip_aggregate(ebi_hip_ctx *ctx, const ebi_hip_col *const *cols, size_t n_cols,
                      ebi_aggregation *aggs, size_t n_aggs, const uint64_t *bitset);

/* ------------------------------------------------------------------ */
/* Serving: continuous batching of concurrent queries over resident     */
/* columns.  No counterpart in the reference (its Decoder is             */
/* single-threaded and synchronous, api/decoder.rs:202-439); each query  */
/* gives the same result as the synchronous call.                        */
/* ------------------------------------------------------------------ */
typedef struct ebi_hip_server ebi_hip_server;
typedef struct ebi_serve_opts {
    uint64_t device_budget; /* HBM bytes for in-flight query scratch (0: 8 GiB) */
    uint64_t pinned_budget; /* pinned host bytes for staging (0: 256 MiB) */
    uint32_t max_inflight;  /* queries on the device at once (0: 256) */
    uint32_t n_streams;     /* HIP streams batches rotate over (0: 4) */
    uint32_t max_batch;     /* queries admitted into one batch (0: 64) */
    uint32_t reserved;
} ebi_serve_opts;

typedef struct ebi_serve_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t rejected;        /* needed more memory than the budgets hold */
    uint64_t batches;         /* batched launches issued */
    uint64_t max_inflight;    /* most queries on the device at once */
    uint64_t device_bytes_peak;
} ebi_serve_stats;

/* Starts the scheduler thread of a server on ctx's device.  The columns
 * queried must stay loaded while their queries are in flight. */
int ebi_hip_serve_create(ebi_hip_ctx *ctx, const ebi_serve_opts *opts, ebi_hip_server **out);
/* Queues n queries (thread-safe); tickets[i] names query i.  Each query's
 * buffers are borrowed until its ticket completes.  Queries are admitted in
 * arrival order as in-flight slots and memory budget free up; every query
 * admitted in one scheduler pass shares one batched decode launch, and
 * passes rotate over the streams so their work overlaps on the device. */
int ebi_hip_serve_submit(ebi_hip_server *srv, const ebi_query *q, size_t n, uint64_t *tickets);
/* Non-blocking: EBI_OK and *status = the query's ebi_status once it has
 * completed, EBI_PENDING before. */
int ebi_hip_serve_poll(ebi_hip_server *srv, uint64_t ticket, int *status);
/* Waits up to timeout_ms (negative: forever).  EBI_OK once completed, with
 * *status and *latency_ms (submit to completion, host clock); the ticket is
 * then released.  EBI_PENDING on timeout. */
int ebi_hip_serve_wait(ebi_hip_server *srv, uint64_t ticket, int64_t timeout_ms, int *status,
                       double *latency_ms);
int ebi_hip_serve_stats(const ebi_hip_server *srv, ebi_serve_stats *st);
/* Drains every queued query, stops the scheduler and frees the server. */
int ebi_hip_serve_destroy(ebi_hip_server *srv);

/* Device allocation helpers for callers without their own allocator. */
int ebi_hip_device_alloc(ebi_hip_ctx *ctx, size_t bytes, void **d_ptr);
int ebi_hip_device_free(ebi_hip_ctx *ctx, void *d_ptr);
int ebi_hip_copy_to_host(ebi_hip_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int ebi_hip_copy_to_device(ebi_hip_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);

/* The same write path on the GPU for BUFF and DeltaSprintz (the two codecs
 * whose chunks are fixed-width packings): d_values are n f64 values in
 * device memory (HBM); the chunk bodies are encoded by kernels and the file
 * comes back in *out (free with ebi_buffer_free), byte-identical to
 * ebi_encode's.  Other schemes return EBI_UNSUPPORTED.  Replaces
 * ChunkWriter::write (encoder.rs:317-405) for those codecs. */
int ebi_hip_encode_device(ebi_hip_ctx *ctx, const double *d_values, size_t n,
                          const ebi_encoder_config *cfg, uint8_t **out, size_t *out_len);
/* Encodes d_values on the GPU as ebi_hip_encode_device and loads the result
 * as a resident column (e.g. a decoded column, changed or appended in HBM,
 * re-encoded without moving its records through the host). */
int ebi_hip_column_encode(ebi_hip_ctx *ctx, const double *d_values, size_t n,
                          const ebi_encoder_config *cfg, ebi_hip_col **out);

/* Zeroes `bytes` (a multiple of 4) of device memory with a fill kernel. */
int ebi_hip_device_zero(ebi_hip_ctx *ctx, void *d_ptr, size_t bytes);

/* Milliseconds of device time of the last kernel sequence issued by the
 * previous call on this ctx (hipEvent pair on the ctx stream). */
double ebi_hip_last_device_ms(const ebi_hip_ctx *ctx);

/* Device time of each phase of the last decode or matmul call on this ctx:
 * for ebi_hip_decode_device / ebi_hip_decode_batch phase 0 = the persistent
 * tile kernel (Uncompressed, RLE, BUFF, DeltaSprintz), phase 2 = the
 * persistent segment kernel (Gorilla, Chimp, Elf, Chimp128), phase 1 = 0 (no
 * kernel); for an eager ebi_hip_batch_run of a mixed batch phase 0 = the one
 * mixed kernel and phase 2 = 0; for the matmul entry points the phases listed
 * there.
 * Instrumentation of this library; no counterpart in the reference. */
double ebi_hip_last_phase_ms(const ebi_hip_ctx *ctx, int phase);

#ifdef __cplusplus
}
#endif

#endif /* EBI_HIP_H */
