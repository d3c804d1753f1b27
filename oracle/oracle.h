/*
 * oracle.h -- CPU restatement of lemolatoon/ebi's chunked decoders and query
 * path.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP library is
 * compared against (tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg).  No product code path may call it.
 *
 * Every function follows the reference file:line cited beside it in
 * oracle.c.  Semantics are those of the reference's release build
 * (wrapping integer arithmetic, saturating float->int casts); inputs that
 * would make the reference panic are reported as ORC_CORRUPT.
 */
#ifndef EBI_ORACLE_H
#define EBI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_UNEXPECTED_END = 1, /* DecoderError::UnexpectedEndOfChunk / IO */
    ORC_CONVERSION = 2,     /* DecoderError::ConversionError */
    ORC_CORRUPT = 3,        /* stream the reference would panic on */
    ORC_UNSUPPORTED = 8,
    ORC_NOMEM = 9,
    ORC_CAPACITY = 10       /* caller buffer too small */
};

/* Same layout as ebi_pred in include/ebi_hip.h. */
typedef struct orc_pred {
    int32_t kind;       /* 0 Eq, 1 Ne, 2 Range */
    int32_t start_kind; /* 0 None, 1 Inclusive, 2 Exclusive */
    int32_t end_kind;
    int32_t reserved;
    double value;
    double start;
    double end;
} orc_pred;

/* File metadata (decoder.rs:131-220). */
int orc_file_records(const uint8_t *file, size_t len, uint64_t *n_records, uint64_t *n_chunks);

/* Decoder::materialize (api/decoder.rs:202-237).  bitset nullable, words =
 * ceil(N/64) over global ids; chunk_id < 0 means all chunks. */
int orc_materialize(const uint8_t *file, size_t len, const uint64_t *bitset, int64_t chunk_id,
                    double *out, size_t cap, size_t *n_out);
/* The same after Decoder::with_precision(precision) (api/decoder.rs:159-161;
 * BUFF decodes with decode_with_precision, chunk_reader/buff/
 * controlled_precision.rs:5-75); precision < 0: full precision. */
int orc_materialize_p(const uint8_t *file, size_t len, const uint64_t *bitset, int64_t chunk_id,
                      int32_t precision, double *out, size_t cap, size_t *n_out);
int orc_filter_materialize_p(const uint8_t *file, size_t len, const orc_pred *pred, const uint64_t *bitset,
                             int64_t chunk_id, int32_t precision, double *out, size_t cap, size_t *n_out);
/* Decoder::filter (api/decoder.rs:242-281); out has ceil(N/64) words. */
int orc_filter(const uint8_t *file, size_t len, const orc_pred *pred, const uint64_t *bitset,
               int64_t chunk_id, uint64_t *out_bitset);
/* Decoder::filter_materialize (api/decoder.rs:283-319). */
int orc_filter_materialize(const uint8_t *file, size_t len, const orc_pred *pred,
                           const uint64_t *bitset, int64_t chunk_id, double *out, size_t cap,
                           size_t *n_out);
/* Decoder::{sum, min, max} (api/decoder.rs:323-439). */
int orc_sum(const uint8_t *file, size_t len, const uint64_t *bitset, int64_t chunk_id, double *out);
int orc_min(const uint8_t *file, size_t len, const uint64_t *bitset, int64_t chunk_id, double *out);
int orc_max(const uint8_t *file, size_t len, const uint64_t *bitset, int64_t chunk_id, double *out);
/* FP64 compensated (Neumaier) sum of the decoded values selected by the
 * bitset and chunk_id: the tolerance reference for GPU sums. */
int orc_sum_compensated(const uint8_t *file, size_t len, const uint64_t *bitset, int64_t chunk_id,
                        double *out);

/* GeneralChunkReader::{distance_squared (dot = 0), dot_product (dot = 1)}
 * on chunk chunk_id (decoder/query.rs:157-192). */
int orc_chunk_fold(const uint8_t *file, size_t len, uint64_t chunk_id, size_t offset, const double *target,
                   size_t tlen, int dot, double *out);
/* Decoder::knn1 (api/decoder.rs:469-533). */
int orc_knn1(const uint8_t *file, size_t len, const double *target, size_t tlen, uint64_t *index, double *dist);
/* Decoder::matmul (api/decoder.rs:606-656); status 3 = PreconditionsNotMet. */
int orc_matmul_dot(const uint8_t *file, size_t len, const double *target, size_t trows, size_t tcols,
                   size_t drows, size_t dcols, double *out);
/* Aggregation::compute_multiple of one aggregation (api/decoder.rs:874-917,
 * decoder/expr.rs:67-123); nodes as ebi_expr_node; kind 0 none, 1 max, 2 min,
 * 3 sum; status 3 = PreconditionsNotMet. */
typedef struct orc_node_ { int32_t kind; uint32_t index; double literal; } orc_node_t;
int orc_aggregate(const uint8_t *const *files, const size_t *lens, size_t n_files, const void *nodes,
                  size_t n_nodes, int kind, const uint64_t *bitset, double *out_scalar, double *sum_comp,
                  double *out_values, size_t cap, size_t *n_out);

/* Helpers exposed for known-answer tests. */
uint64_t orc_zigzag(int64_t v);
int64_t orc_unzigzag(uint64_t v);
int64_t orc_buff_fixed(double v, int32_t frac);
double orc_elf_round_up(double v, uint32_t alpha);
int32_t orc_elf_get_sp(double v);

#ifdef __cplusplus
}
#endif
#endif
