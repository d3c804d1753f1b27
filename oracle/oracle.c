/*
 * oracle.c -- CPU restatement of the reference chunk decoders and query
 * operators.  TEST INFRASTRUCTURE ONLY (see oracle.h): it checks the HIP
 * path and serves as the timed CPU baseline; it is never the product path.
 *
 * It deliberately keeps the reference's structure: bit-serial MSB readers,
 * iterator-style decoders that stop at the NaN terminator, BUFF queries that
 * walk byte planes MSB-first with candidate sets, and the Decoder's per-chunk
 * loops with bitmask pruning.  Compile with -ffp-contract=off.
 *
 * Citations are relative to the lemolatoon/ebi repository.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NAN_BITS 0x7ff8000000000000ull

static uint64_t rd64(const uint8_t *p) { uint64_t v; memcpy(&v, p, 8); return v; }
static uint32_t rd32(const uint8_t *p) { uint32_t v; memcpy(&v, p, 4); return v; }
static double bits2d(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
static uint64_t d2bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

/* Rust `as i64` / `as i32` / `as u32` from f64: saturating, NaN -> 0. */
static int64_t sat_i64(double d) {
    if (isnan(d)) return 0;
    if (d >= 9223372036854775808.0) return INT64_MAX;
    if (d < -9223372036854775808.0) return INT64_MIN;
    return (int64_t)d;
}
static int32_t sat_i32(double d) {
    if (isnan(d)) return 0;
    if (d >= 2147483647.0) return INT32_MAX;
    if (d <= -2147483648.0) return INT32_MIN;
    return (int32_t)d;
}

/* ------------------------------------------------------------------ */
/* File parse: decoder.rs:131-220, format/deserialize.rs:117-157,      */
/* GeneralChunkHandle geometry decoder.rs:283-352.                     */
/* ------------------------------------------------------------------ */
typedef struct {
    int scheme;
    uint64_t n_records, n_chunks;
    uint64_t *phys, *size, *logi, *nrec;
} orc_file;

static void file_free(orc_file *f) {
    free(f->phys); free(f->size); free(f->logi); free(f->nrec);
    memset(f, 0, sizeof *f);
}

static int file_parse(const uint8_t *b, size_t len, orc_file *f) {
    memset(f, 0, sizeof *f);
    if (len < 30) return ORC_UNEXPECTED_END;
    if (b[4] != 0) return ORC_CONVERSION;      /* FieldType::F64 only */
    if (b[5] > 12) return ORC_CONVERSION;      /* CompressionScheme::try_from_primitive */
    f->scheme = b[5];
    uint64_t fo = rd64(b + 16);
    if (fo < 30 || fo > len || len - fo < 16) return ORC_UNEXPECTED_END;
    f->n_records = rd64(b + fo);
    f->n_chunks = rd64(b + fo + 8);
    uint64_t pos = fo + 16;
    if (f->n_chunks > (len - pos) / 16) return ORC_UNEXPECTED_END;
    size_t nc = (size_t)f->n_chunks;
    f->phys = calloc(nc + 1, 8); f->size = calloc(nc + 1, 8);
    f->logi = calloc(nc + 1, 8); f->nrec = calloc(nc + 1, 8);
    if (!f->phys || !f->size || !f->logi || !f->nrec) { file_free(f); return ORC_NOMEM; }
    for (size_t c = 0; c < nc; ++c) {
        f->phys[c] = rd64(b + pos);
        f->logi[c] = rd64(b + pos + 8);
        pos += 16;
    }
    if (pos + 1 > len) { file_free(f); return ORC_UNEXPECTED_END; }
    size_t cfg = b[pos];
    if (pos + 1 + cfg + 164 > len) { file_free(f); return ORC_UNEXPECTED_END; }
    for (size_t c = 0; c < nc; ++c) {
        uint64_t np = c + 1 < nc ? f->phys[c + 1] : fo;
        uint64_t nl = c + 1 < nc ? f->logi[c + 1] : f->n_records;
        if (f->phys[c] < 30 || np < f->phys[c] || np > fo || nl < f->logi[c]) {
            file_free(f);
            return ORC_UNEXPECTED_END;
        }
        f->size[c] = np - f->phys[c];
        f->nrec[c] = nl - f->logi[c];
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* MSB-first bit reader, io/bit_read.rs:37-368                         */
/* ------------------------------------------------------------------ */
typedef struct {
    const uint8_t *buf;
    size_t len, byte_pos;
    unsigned bit_pos;
} msb_t;

static int msb_bit(msb_t *r, int *bit) {            /* read_bit :291-299 */
    if (r->byte_pos >= r->len) return -1;
    *bit = (r->buf[r->byte_pos] >> (7 - r->bit_pos)) & 1;
    if (++r->bit_pos == 8) { r->bit_pos = 0; r->byte_pos++; }
    return 0;
}
static uint8_t msb_byte(msb_t *r) {                 /* read_byte_unchecked :143-159 */
    uint8_t hi = r->buf[r->byte_pos++];
    if (r->bit_pos == 0) return hi;
    uint8_t lo = r->byte_pos < r->len ? (uint8_t)(r->buf[r->byte_pos] >> (8 - r->bit_pos)) : 0;
    return (uint8_t)(hi << r->bit_pos) | lo;
}
/* read_bits :293-318: whole bytes first, then the remaining bits one by one. */
static int msb_bits(msb_t *r, unsigned n, uint64_t *out) {
    if ((r->len - r->byte_pos) * 8 - r->bit_pos < n) return -1;
    uint64_t v = 0;
    unsigned rem = n;
    for (; rem >= 8; rem -= 8) v = (v << 8) | msb_byte(r);
    for (; rem; --rem) {
        int b = 0;
        msb_bit(r, &b);
        v = (v << 1) | (uint64_t)b;
    }
    *out = v;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Growable value vector                                               */
/* ------------------------------------------------------------------ */
typedef struct { double *v; size_t n, cap; } dvec;
static int dpush(dvec *d, double x) {
    if (d->n == d->cap) {
        size_t nc = d->cap ? d->cap * 2 : 64;
        double *p = realloc(d->v, nc * sizeof(double));
        if (!p) return ORC_NOMEM;
        d->v = p; d->cap = nc;
    }
    d->v[d->n++] = x;
    return ORC_OK;
}

static const uint32_t LEAD_REPR[8] = {0, 8, 12, 16, 18, 20, 22, 24};

#define RD(r, n, out) do { if (msb_bits((r), (n), (out))) return ORC_UNEXPECTED_END; } while (0)
#define RB(r, out) do { if (msb_bit((r), (out))) return ORC_UNEXPECTED_END; } while (0)

/* Gorilla, decoder/chunk_reader/gorilla.rs:77-157 (via general_xor.rs:118-133). */
static int dec_gorilla(msb_t *r, dvec *o) {
    uint64_t v;
    uint8_t L = 0, T = 0;
    RD(r, 64, &v);
    for (;;) {
        if (v == NAN_BITS) return ORC_OK;
        if (dpush(o, bits2d(v))) return ORC_NOMEM;
        int c, z;
        RB(r, &c);
        if (!c) continue;
        RB(r, &z);
        if (z) {
            uint64_t l, s;
            RD(r, 5, &l);
            RD(r, 6, &s);
            if (l + s + 1 > 64) return ORC_CORRUPT; /* u8 underflow panics in the reference */
            L = (uint8_t)l;
            T = (uint8_t)(64 - L - (s + 1));
        }
        unsigned size = 64u - L - T;
        uint64_t bits;
        RD(r, size, &bits);
        v ^= T >= 64 ? 0 : bits << T;
    }
}

/* ChimpDecoder::decompress_float, decoder/chunk_reader/chimp.rs:59-125:
 * the first value is 64 raw bits, then a 2-bit flag per value; *has = 0
 * once the value is the NaN end marker. */
typedef struct { uint64_t v; uint32_t lead; int first, end; } chimp_t;
static int chimp_next(chimp_t *s, msb_t *r, uint64_t *out, int *has) {
    *has = 0;
    if (s->end) return ORC_OK;
    uint64_t x, flag, t;
    if (s->first) {
        s->first = 0;
        RD(r, 64, &s->v);
    } else {
        RD(r, 2, &flag);
        switch (flag) {
        case 3:
            RD(r, 3, &t);
            s->lead = LEAD_REPR[t];
            RD(r, 64 - s->lead, &x);
            s->v ^= x;
            break;
        case 2:
            if (s->lead > 64) return ORC_CORRUPT;
            RD(r, 64 - s->lead, &x);
            s->v ^= x;
            break;
        case 1: {
            uint64_t sig;
            RD(r, 3, &t);
            s->lead = LEAD_REPR[t];
            RD(r, 6, &sig);
            if (sig == 0) sig = 64;
            if (sig + s->lead > 64) return ORC_CORRUPT;
            uint32_t tz = 64 - (uint32_t)sig - s->lead;
            RD(r, 64 - s->lead - tz, &x);
            s->v ^= tz >= 64 ? 0 : x << tz;
            break;
        }
        default:
            break;
        }
    }
    if (s->v == NAN_BITS) { s->end = 1; return ORC_OK; }
    *out = s->v; *has = 1;
    return ORC_OK;
}

/* Chimp, decoder/chunk_reader/chimp.rs:59-125 via general_xor.rs:118-133. */
static int dec_chimp(msb_t *r, dvec *o) {
    chimp_t s = {0, UINT32_MAX, 1, 0};
    for (;;) {
        uint64_t v;
        int has, rc;
        if ((rc = chimp_next(&s, r, &v, &has))) return rc;
        if (!has) return ORC_OK;
        if (dpush(o, bits2d(v))) return ORC_NOMEM;
    }
}

/* Chimp128, decoder/chunk_reader/chimp_n.rs:173-285. */
static int dec_chimp128(msb_t *r, dvec *o) {
    uint64_t ring[128] = {0};
    unsigned cur = 0;
    uint32_t lead = UINT32_MAX;
    uint64_t v, x, flag, t;
    RD(r, 64, &v);
    ring[cur] = v;
    if (v == NAN_BITS) return ORC_OK;
    if (dpush(o, bits2d(v))) return ORC_NOMEM;
    for (;;) {
        RD(r, 2, &flag);
        if (flag == 0) {
            RD(r, 7, &t);
            v = ring[t];
            cur = (cur + 1) & 127;
            ring[cur] = v;
            if (dpush(o, bits2d(v))) return ORC_NOMEM; /* no NaN check on this path */
            continue;
        }
        if (flag == 3) {
            RD(r, 3, &t);
            lead = LEAD_REPR[t];
            RD(r, 64 - lead, &x);
            v ^= x;
        } else if (flag == 2) {
            if (lead > 64) return ORC_CORRUPT;
            RD(r, 64 - lead, &x);
            v ^= x;
        } else {
            RD(r, 16, &t); /* INITIAL_FILL = log2(128) + 9 */
            uint64_t idx = (t >> 9) & 127;
            lead = LEAD_REPR[(t >> 6) & 7];
            uint64_t sig = t & 63;
            v = ring[idx];
            if (sig == 0) sig = 64;
            if (sig + lead > 64) return ORC_CORRUPT;
            uint32_t tz = 64 - (uint32_t)sig - lead;
            RD(r, 64 - lead - tz, &x);
            v ^= tz >= 64 ? 0 : x << tz;
        }
        if (v == NAN_BITS) return ORC_OK;
        cur = (cur + 1) & 127;
        ring[cur] = v;
        if (dpush(o, bits2d(v))) return ORC_NOMEM;
    }
}

/* ---- Elf math, compression_common/elf.rs ---- */
static double P10[400], N10[400];
static int p10_ready = 0;
static void p10_init(void) {
    if (p10_ready) return;
    char buf[32];
    for (int i = 0; i < 400; ++i) {
        snprintf(buf, sizeof buf, "1.0E%d", i);   /* get_10i_p :95-101 */
        P10[i] = strtod(buf, NULL);
        snprintf(buf, sizeof buf, "1.0E-%d", i);  /* get_10i_n :103-109 */
        N10[i] = strtod(buf, NULL);
    }
    p10_ready = 1;
}
static double get_10i_p(uint32_t i) { p10_init(); return i < 400 ? P10[i] : INFINITY; }
static double get_10i_n(uint32_t i) { p10_init(); return i < 400 ? N10[i] : 0.0; }

int32_t orc_elf_get_sp(double v) {                /* get_sp :111-130 */
    static const double g1[10] = {1, 10, 100, 1000, 10000, 100000, 1000000, 10000000,
                                   100000000, 1000000000};
    static const double l1[11] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001,
                                  0.0000001, 0.00000001, 0.000000001, 0.0000000001};
    if (v >= 1.0) {
        for (int i = 0; i < 9; ++i)
            if (v < g1[i + 1]) return i;
    } else {
        for (int i = 1; i < 11; ++i)
            if (v >= l1[i]) return -i;
    }
    return sat_i32(floor(log10(v)));
}

double orc_elf_round_up(double v, uint32_t alpha) { /* round_up :51-58 */
    double scale = get_10i_p(alpha);
    if (v < 0.0) return floor(v * scale) / scale;
    return ceil(v * scale) / scale;
}

typedef struct {
    uint64_t stored;
    uint32_t L, T;
    int first, end;
} elfxor_t;

/* ElfXorDecoder, decoder/chunk_reader/elf.rs:105-258.  *has = 0 on end. */
static int elfxor_next(elfxor_t *s, msb_t *r, uint64_t *out, int *has) {
    *has = 0;
    if (s->end) return ORC_OK;
    uint64_t t, x;
    if (s->first) {
        s->first = 0;
        RD(r, 7, &t);
        if (t < 64) {
            RD(r, 63 - (unsigned)t, &x);
            s->stored = ((x << 1) | 1) << t;
        } else {
            s->stored = 0;
        }
        if (s->stored == NAN_BITS) { s->end = 1; return ORC_OK; }
        *out = s->stored; *has = 1;
        return ORC_OK;
    }
    uint64_t flag;
    RD(r, 2, &flag);
    uint64_t value;
    if (flag == 3) {
        RD(r, 9, &t);
        s->L = LEAD_REPR[t >> 6];
        uint32_t center = (uint32_t)(t & 63);
        if (center == 0) center = 64;
        if (s->L + center > 64) return ORC_CORRUPT;
        s->T = 64 - s->L - center;
        RD(r, center - 1, &x);
        value = (((x << 1) | 1) << s->T) ^ s->stored;
        if (value == NAN_BITS) { s->end = 1; return ORC_OK; }
        s->stored = value;
    } else if (flag == 2) {
        RD(r, 7, &t);
        s->L = LEAD_REPR[t >> 4];
        uint32_t center = (uint32_t)(t & 15);
        if (center == 0) center = 16;
        if (s->L + center > 64) return ORC_CORRUPT;
        s->T = 64 - s->L - center;
        if (center == 1) {
            value = 1ull << s->T;
        } else {
            RD(r, center - 1, &x);
            value = ((x << 1) | 1) << s->T;
        }
        value ^= s->stored;
        if (value == NAN_BITS) { s->end = 1; return ORC_OK; }
        s->stored = value;
    } else if (flag == 0) {
        if ((uint64_t)s->L + s->T > 64) return ORC_CORRUPT;
        uint32_t center = 64 - s->L - s->T;
        RD(r, center, &x);
        value = (s->T >= 64 ? 0 : x << s->T) ^ s->stored;
        if (value == NAN_BITS) { s->end = 1; return ORC_OK; }
        s->stored = value;
    }
    *out = s->stored; *has = 1;
    return ORC_OK;
}

/* ElfDecoderWrapper, decoder/chunk_reader/elf.rs:46-93, over the ElfXor
 * decoder (Elf, scheme 7) or the Chimp decoder (ElfOnChimp, scheme 6,
 * elf.rs:27). */
static int dec_elf(msb_t *r, dvec *o, int on_chimp) {
    elfxor_t s = {0, UINT32_MAX, UINT32_MAX, 1, 0};
    chimp_t c = {0, UINT32_MAX, 1, 0};
    int32_t last_beta_star = INT32_MAX;
    for (;;) {
        int b, has, rc;
        uint64_t vp;
        RB(r, &b);
        int recover;
        if (!b) {
            recover = 1;                                  /* case 0 */
        } else {
            RB(r, &b);
            if (!b) {
                recover = 0;                              /* case 10 */
            } else {
                uint64_t bs;
                RD(r, 4, &bs);
                last_beta_star = (int32_t)bs;             /* case 11 */
                recover = 1;
            }
        }
        rc = on_chimp ? chimp_next(&c, r, &vp, &has) : elfxor_next(&s, r, &vp, &has);
        if (rc) return rc;
        if (!has) return ORC_OK;
        double v = bits2d(vp);
        if (recover) {                                    /* recover_v_by_beta_star :66-81 */
            int32_t sp = orc_elf_get_sp(fabs(v));
            if (last_beta_star == 0) {
                double w = get_10i_n(0u - (uint32_t)sp - 1u);  /* (-sp - 1) as u32, wrapping */
                v = v < 0.0 ? -w : w;
            } else {
                uint32_t alpha = (uint32_t)last_beta_star - (uint32_t)sp - 1u; /* wrapping i32 */
                v = orc_elf_round_up(v, alpha);
            }
        }
        if (dpush(o, v)) return ORC_NOMEM;
    }
}

/* Sprintz decode iterator, decoder/chunk_reader/sprintz.rs:448-504. */
typedef struct { int64_t q0; uint32_t scale; unsigned W; } spz_hdr;
static int spz_header(msb_t *r, spz_hdr *h) {
    uint64_t t;
    RD(r, 64, &t); h->q0 = (int64_t)t;
    RD(r, 32, &t); h->scale = (uint32_t)t;
    RD(r, 8, &t); h->W = (unsigned)t;
    return ORC_OK;
}
int64_t orc_unzigzag(uint64_t v) { return (int64_t)(v >> 1) ^ -(int64_t)(v & 1); } /* :519-521 */
uint64_t orc_zigzag(int64_t v) {                                                       /* compressor/sprintz.rs:232-235 */
    return (uint64_t)(((uint64_t)v << 1) ^ (uint64_t)(v >> 63));
}

/* Walk the quantized values q_0..q_{n-1} (n = footer record count). */
static int spz_walk(msb_t *r, uint64_t n, spz_hdr *h, int64_t *q) {
    int rc = spz_header(r, h);
    if (rc) return rc;
    int64_t prev = h->q0;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t z = 0;
        if (h->W) RD(r, h->W, &z);
        prev = (int64_t)((uint64_t)prev + (uint64_t)orc_unzigzag(z));
        q[i] = prev;
    }
    return ORC_OK;
}

/* ---- BUFF: compression_common/buff.rs and decoder/chunk_reader/buff/ ---- */
typedef struct {
    int64_t base;
    uint32_t n, B, frac;
    const uint8_t *planes; /* floor(B/8) planes of n bytes */
    const uint8_t *tail;   /* LSB-first (B % 8)-bit fields, or B-bit if B < 8 */
    unsigned n_planes, r;
    double scale;
} buff_t;

static int buff_open(const uint8_t *p, size_t size, buff_t *b) {
    if (size < 20) return ORC_CORRUPT;   /* read_u32().unwrap() panics */
    uint64_t lo = rd32(p), hi = rd32(p + 4);
    b->base = (int64_t)(lo | (hi << 32));
    b->n = rd32(p + 8);
    b->B = rd32(p + 12);
    b->frac = rd32(p + 16);
    if (b->B > 64) return ORC_CORRUPT;
    b->scale = pow(2.0, (double)b->frac); /* 2.0f64.powi(frac): exact */
    b->n_planes = b->B >= 8 ? b->B / 8 : 0;
    b->r = b->B >= 8 ? b->B % 8 : b->B;
    b->planes = p + 20;
    b->tail = p + 20 + (size_t)b->n_planes * b->n;
    size_t need = 20 + (size_t)b->n_planes * b->n + ((size_t)b->n * b->r + 7) / 8;
    if (need > size) return ORC_CORRUPT;
    return ORC_OK;
}
/* BitPack::read_bits, LSB first, a byte's worth of bits at a time
 * (compression_common/buff.rs:206-245); fields are under 8 bits wide. */
static uint32_t buff_tail(const buff_t *b, uint32_t i) {
    uint64_t bit = (uint64_t)i * b->r;
    const uint8_t *q = b->tail + (bit >> 3);
    unsigned at = (unsigned)(bit & 7), left = 8 - at;
    uint32_t v = (uint32_t)(q[0] >> at);
    if (b->r > left) v |= (uint32_t)q[1] << left;
    return v & ((1u << b->r) - 1);
}
static uint8_t flip(uint8_t x) { return x ^ 0x80; } /* buff.rs:2-5 */
static uint8_t plane_byte(const buff_t *b, unsigned pl, uint32_t i) {
    return flip(b->planes[(size_t)pl * b->n + i]);
}
static uint64_t buff_delta(const buff_t *b, uint32_t i) {
    uint64_t d = 0;
    for (unsigned pl = 0; pl < b->n_planes; ++pl) d = (d << 8) | plane_byte(b, pl, i);
    if (b->r) d = (d << b->r) | buff_tail(b, i);
    return d;
}
/* buff_simd256_decode, decoder/chunk_reader/buff/decode.rs:5-70: plane by
 * plane into a fixed-point vector, then the tail fields, then the floats. */
static int dec_buff(const uint8_t *p, size_t size, dvec *o) {
    buff_t b;
    int rc = buff_open(p, size, &b);
    if (rc) return rc;
    uint64_t *fx = calloc(b.n ? b.n : 1, sizeof(uint64_t));
    if (!fx) return ORC_NOMEM;
    for (unsigned pl = 0; pl < b.n_planes; ++pl) {
        const uint8_t *x = b.planes + (size_t)pl * b.n;
        for (uint32_t i = 0; i < b.n; ++i) fx[i] = (fx[i] << 8) | flip(x[i]);
    }
    if (b.r)
        for (uint32_t i = 0; i < b.n; ++i) fx[i] = (fx[i] << b.r) | buff_tail(&b, i);
    for (uint32_t i = 0; i < b.n; ++i)
        if (dpush(o, (double)(int64_t)((uint64_t)b.base + fx[i]) / b.scale)) { free(fx); return ORC_NOMEM; }
    free(fx);
    return ORC_OK;
}

/* decode_with_precision, decoder/chunk_reader/buff/controlled_precision.rs:5-75
 * (Decoder::with_precision, api/decoder.rs:159-161): the first byte plane
 * always, further planes while the bits below them are still needed, the
 * LSB tail only if every plane was read and its bits are needed. */
static const uint32_t PRECISION_MAP[13] = {0, 5, 8, 11, 15, 18, 21, 25, 28, 31, 35, 38, 50}; /* buff.rs:36 */
static int dec_buff_precision(const uint8_t *p, size_t size, uint32_t precision, dvec *o) {
    buff_t b;
    int rc = buff_open(p, size, &b);
    if (rc) return rc;
    if (precision > 12) return ORC_CORRUPT;       /* PRECISION_MAP[precision] panics */
    uint32_t keep = PRECISION_MAP[precision];
    uint32_t unnecessary = b.frac > keep ? b.frac - keep : 0; /* saturating_sub */
    uint32_t remaining = b.B;
    if (remaining < 8) return dec_buff(p, size, o); /* B = 0: base; 1..7: every tail bit */
    uint64_t *fixed = malloc((b.n ? b.n : 1) * sizeof(uint64_t));
    if (!fixed) return ORC_NOMEM;
    remaining -= 8;
    unsigned pl = 0;
    for (uint32_t i = 0; i < b.n; ++i) fixed[i] = (uint64_t)plane_byte(&b, pl, i) << remaining;
    while (remaining >= 8 && remaining > unnecessary) {
        remaining -= 8;
        ++pl;
        for (uint32_t i = 0; i < b.n; ++i) fixed[i] |= (uint64_t)plane_byte(&b, pl, i) << remaining;
    }
    int tail = !(remaining == 0 || remaining <= unnecessary); /* then every plane was read */
    for (uint32_t i = 0; i < b.n; ++i) {
        uint64_t d = fixed[i] | (tail ? buff_tail(&b, i) : 0);
        if (dpush(o, (double)(int64_t)((uint64_t)b.base + d) / b.scale)) { free(fixed); return ORC_NOMEM; }
    }
    free(fixed);
    return ORC_OK;
}

int64_t orc_buff_fixed(double v, int32_t frac) { /* compression_common/buff.rs:10-52 */
    uint64_t bits = d2bits(v);
    int32_t exp = (int32_t)((bits >> 52) & 0x7ff) - 1023;
    if (exp < -frac) return 0;
    int32_t sh = 63 - exp - frac;
    uint64_t fixed = sh >= 64 ? 0 : ((bits << 11) | (1ull << 63)) >> ((uint64_t)(int64_t)sh & 63);
    if (bits >> 63) fixed = ~(fixed - 1);
    return (int64_t)fixed;
}

/* The first plane of a filter without a bitmask, 32 records per step as
 * comp_simd (filter_cmp.rs:484-543) and the equal filter's SIMD form do:
 * signed byte compares of the stored (flipped) bytes with flip(pred) order
 * the unflipped bytes as unsigned.  op 0: equal -> chk; 1: greater -> res,
 * equal -> chk; 2: less -> res, equal -> chk.  Returns the size of chk. */
#if defined(__x86_64__)
#include <immintrin.h>
__attribute__((target("avx2")))
static uint32_t plane_first_avx2(const uint8_t *x, uint32_t n, uint8_t ps, int op,
                                 uint8_t *res, uint32_t *chk) {
    const __m256i pw = _mm256_set1_epi8((char)flip(ps));
    uint32_t m = 0, i = 0;
    for (; i + 32 <= n; i += 32) {
        __m256i w = _mm256_loadu_si256((const __m256i *)(x + i));
        uint32_t eq = (uint32_t)_mm256_movemask_epi8(_mm256_cmpeq_epi8(w, pw));
        if (op) {
            uint32_t q = (uint32_t)_mm256_movemask_epi8(op == 1 ? _mm256_cmpgt_epi8(w, pw)
                                                                : _mm256_cmpgt_epi8(pw, w));
            for (; q; q &= q - 1) res[i + (uint32_t)__builtin_ctz(q)] = 1;
        }
        for (; eq; eq &= eq - 1) chk[m++] = i + (uint32_t)__builtin_ctz(eq);
    }
    for (; i < n; ++i) {
        uint8_t v = flip(x[i]);
        if (v == ps) chk[m++] = i;
        else if (op && ((op == 1) == (v > ps))) res[i] = 1;
    }
    return m;
}
#endif
static uint32_t plane_first(const uint8_t *x, uint32_t n, uint8_t ps, int op, uint8_t *res, uint32_t *chk) {
#if defined(__x86_64__)
    static int avx2 = -1;
    if (avx2 < 0) { __builtin_cpu_init(); avx2 = __builtin_cpu_supports("avx2") ? 1 : 0; }
    if (avx2) return plane_first_avx2(x, n, ps, op, res, chk);
#endif
    uint32_t m = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint8_t v = flip(x[i]);
        if (v == ps) chk[m++] = i;
        else if (op && ((op == 1) == (v > ps))) res[i] = 1;
    }
    return m;
}

/* buff_simd256_cmp_filter::<IS_GREATER, INCLUSIVE>, filter_cmp.rs:298-451.
 * mask (nullable) = the bitmask restricted to the chunk; res[i] out.  The
 * records still undecided after each plane are kept as an ascending index
 * list (the reference's to_check bitmap), so later planes and the tail
 * visit only those. */
static int buff_cmp(const buff_t *b, double pred, int greater, int inclusive,
                    const uint8_t *mask, uint8_t *res) {
    uint32_t n = b->n;
    memset(res, 0, n);
    int64_t fp = orc_buff_fixed(pred, (int32_t)b->frac);
    if (fp < b->base) {                                     /* :328-330 */
        if (greater) memset(res, 1, n);
        return ORC_OK;
    }
    if (b->B == 0) {                                        /* :331-339 */
        int all = greater ? (inclusive ? b->base >= fp : b->base > fp)
                          : (inclusive ? b->base <= fp : b->base < fp);
        if (all) memset(res, 1, n);
        return ORC_OK;
    }
    uint64_t dp = (uint64_t)fp - (uint64_t)b->base;
    unsigned pbits = dp == 0 ? 0 : 64 - (unsigned)__builtin_clzll(dp);
    if (b->B < pbits) {                                     /* :342-351 */
        if (!greater) memset(res, 1, n);
        return ORC_OK;
    }
    uint32_t *chk = malloc((n ? n : 1) * sizeof(uint32_t));
    if (!chk) return ORC_NOMEM;
    uint32_t nchk = 0;
    unsigned rem = b->B, pl = 0;
    if (!mask && rem >= 8) {                                /* :365-371, no bitmask */
        rem -= 8;
        nchk = plane_first(b->planes, n, (uint8_t)(dp >> rem), greater ? 1 : 2, res, chk);
        pl = 1;
        if (!nchk) { free(chk); return ORC_OK; }
    } else {
        for (uint32_t i = 0; i < n; ++i) if (!mask || mask[i]) chk[nchk++] = i;
    }
    while (rem >= 8) {                                      /* :357-385 */
        rem -= 8;
        uint8_t ps = (uint8_t)(dp >> rem);
        const uint8_t *x = b->planes + (size_t)pl * n;
        uint32_t m = 0;
        for (uint32_t k = 0; k < nchk; ++k) {
            uint32_t i = chk[k];
            uint8_t v = flip(x[i]);
            if (v == ps) chk[m++] = i;
            else if (greater == (v > ps)) res[i] = 1;
        }
        nchk = m;
        ++pl;
        if (!nchk) { free(chk); return ORC_OK; }
    }
    if (rem > 0) {                                          /* :388-444 */
        uint32_t ps = (uint32_t)(dp & ((1u << rem) - 1));
        for (uint32_t k = 0; k < nchk; ++k) {
            uint32_t i = chk[k], v = buff_tail(b, i);
            if (v > ps) { if (greater) res[i] = 1; }
            else if (v == ps) { if (inclusive) res[i] = 1; }
            else { if (!greater) res[i] = 1; }
        }
    } else if (inclusive) {                                 /* :445-449 */
        for (uint32_t k = 0; k < nchk; ++k) res[chk[k]] = 1;
    }
    free(chk);
    return ORC_OK;
}

/* buff_simd256_equal_filter, filter_eq.rs:73-201; the candidates after each
 * plane are an ascending index list (the reference's result_bitmask). */
static int buff_eq(const buff_t *b, double pred, const uint8_t *mask, uint8_t *res) {
    uint32_t n = b->n;
    memset(res, 0, n);
    int64_t fp = orc_buff_fixed(pred, (int32_t)b->frac);
    if (fp < b->base) return ORC_OK;
    if (b->B == 0) {
        if (fp == b->base) memset(res, 1, n);
        return ORC_OK;
    }
    uint64_t dp = (uint64_t)fp - (uint64_t)b->base;
    unsigned rem = b->B, pl = 0;
    if (!mask && rem < 8) {                 /* no bitmask, no plane: :181-191 */
        uint8_t ps = (uint8_t)dp;           /* `delta_fixed_pred as u8`, unmasked */
        for (uint32_t i = 0; i < n; ++i) res[i] = buff_tail(b, i) == ps;
        return ORC_OK;
    }
    uint32_t *chk = malloc((n ? n : 1) * sizeof(uint32_t));
    if (!chk) return ORC_NOMEM;
    uint32_t nchk = 0;
    if (!mask) {                            /* first plane without a bitmask */
        rem -= 8;
        nchk = plane_first(b->planes, n, (uint8_t)(dp >> rem), 0, res, chk);
        pl = 1;
        if (!nchk) { free(chk); return ORC_OK; }
    } else {
        for (uint32_t i = 0; i < n; ++i) if (mask[i]) chk[nchk++] = i;
    }
    while (rem >= 8) {
        rem -= 8;
        uint8_t ps = (uint8_t)(dp >> rem);  /* truncating cast: quirk, no width check */
        const uint8_t *x = b->planes + (size_t)pl * n;
        uint32_t m = 0;
        for (uint32_t k = 0; k < nchk; ++k)
            if (flip(x[chk[k]]) == ps) chk[m++] = chk[k];
        nchk = m;
        ++pl;
        if (!nchk) { free(chk); return ORC_OK; }
    }
    if (rem > 0) {
        uint32_t ps = (uint32_t)(dp & ((1u << rem) - 1));
        for (uint32_t k = 0; k < nchk; ++k)
            if (buff_tail(b, chk[k]) == ps) res[chk[k]] = 1;
    } else {
        for (uint32_t k = 0; k < nchk; ++k) res[chk[k]] = 1;
    }
    free(chk);
    return ORC_OK;
}

/* BUFFReader::filter, decoder/chunk_reader/buff.rs:128-228. */
static int buff_filter(const buff_t *b, const orc_pred *p, const uint8_t *mask, uint8_t *res) {
    uint32_t n = b->n;
    int rc;
    if (p->kind == 0 || p->kind == 1) {
        if ((rc = buff_eq(b, p->value, mask, res))) return rc;
        if (p->kind == 1)
            for (uint32_t i = 0; i < n; ++i) res[i] = (mask ? mask[i] : 1) && !res[i];
        return ORC_OK;
    }
    uint8_t *left = NULL, *right = NULL;
    if (p->start_kind != 0) {
        left = malloc(n ? n : 1);
        if (!left) return ORC_NOMEM;
        if ((rc = buff_cmp(b, p->start, 1, p->start_kind == 1, mask, left))) { free(left); return rc; }
    }
    if (p->end_kind != 0) {
        right = malloc(n ? n : 1);
        if (!right) { free(left); return ORC_NOMEM; }
        if ((rc = buff_cmp(b, p->end, 0, p->end_kind == 1, mask, right))) { free(left); free(right); return rc; }
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (!left && !right) res[i] = mask ? mask[i] : 1;
        else if (!left) res[i] = right[i];
        else if (!right) res[i] = left[i];
        else res[i] = left[i] && right[i];
    }
    free(left);
    free(right);
    return ORC_OK;
}

/* sum_with_bitmask, buff/sum.rs:60-119. */
static double buff_sum(const buff_t *b, const uint8_t *mask) {
    if (b->B == 0) return ((double)b->base / b->scale) * (double)b->n;
    uint64_t sum = 0, cnt = 0;
    for (unsigned pl = 0; pl < b->n_planes; ++pl) {
        uint64_t s = 0;
        for (uint32_t i = 0; i < b->n; ++i)
            if (!mask || mask[i]) s += plane_byte(b, pl, i);
        sum += s << ((b->n_planes - 1 - pl) * 8 + b->r);
    }
    if (b->r) {
        uint64_t s = 0;
        for (uint32_t i = 0; i < b->n; ++i)
            if (!mask || mask[i]) s += buff_tail(b, i);
        sum += s;
    }
    if (mask) { for (uint32_t i = 0; i < b->n; ++i) cnt += mask[i] != 0; }
    else cnt = b->n;
    return (double)(int64_t)(sum + (uint64_t)b->base * cnt) / b->scale;
}

/* max_with_bitmask / max_without_bitmask / min_without_bitmask, buff/min_max.rs.
 * Plane-wise descent over the candidate set, then reassembly of the winner. */
static double buff_minmax(const buff_t *b, const uint8_t *mask, int is_min) {
    if (b->B == 0) return (double)b->base / b->scale;
    if (is_min && !mask) return (double)b->base / b->scale;  /* min_without_bitmask :316-341 */
    uint32_t n = b->n;
    uint8_t *check = malloc(n ? n : 1);
    for (uint32_t i = 0; i < n; ++i) check[i] = mask ? mask[i] : 1;
    unsigned rem = b->B, pl = 0;
    while (rem > 0) {
        unsigned best = is_min ? 0xffffffffu : 0;
        int use_tail = rem < 8;
        for (uint32_t i = 0; i < n; ++i) {
            if (!check[i]) continue;
            unsigned s = use_tail ? buff_tail(b, i) : plane_byte(b, pl, i);
            if (is_min ? s < best : s > best) best = s;
        }
        for (uint32_t i = 0; i < n; ++i) {
            if (!check[i]) continue;
            unsigned s = use_tail ? buff_tail(b, i) : plane_byte(b, pl, i);
            if (s != best) check[i] = 0;
        }
        if (use_tail) break;
        rem -= 8;
        ++pl;
    }
    uint32_t win = 0;
    while (win < n && !check[win]) ++win;
    free(check);
    uint64_t d = win < n ? buff_delta(b, win) : 0;
    return (double)(int64_t)((uint64_t)b->base + d) / b->scale;
}

/* ------------------------------------------------------------------ */
/* Predicate::eval, decoder/query.rs:374-446                           */
/* ------------------------------------------------------------------ */
static int pred_eval_f(const orc_pred *p, double v) {
    if (p->kind == 0) return v == p->value;
    if (p->kind == 1) return v != p->value;
    int s = p->start_kind == 0 ? 1 : p->start_kind == 1 ? v >= p->start : v > p->start;
    int e = p->end_kind == 0 ? 1 : p->end_kind == 1 ? v <= p->end : v < p->end;
    return s && e;
}
static int pred_eval_q(const orc_pred *p, double scale, int64_t q) {
    /* predicate.map(|v| (v * scale).round() as i64), sprintz.rs:116 */
    if (p->kind == 0) return q == sat_i64(round(p->value * scale));
    if (p->kind == 1) return q != sat_i64(round(p->value * scale));
    int s = 1, e = 1;
    if (p->start_kind) {
        int64_t a = sat_i64(round(p->start * scale));
        s = p->start_kind == 1 ? q >= a : q > a;
    }
    if (p->end_kind) {
        int64_t a = sat_i64(round(p->end * scale));
        e = p->end_kind == 1 ? q <= a : q < a;
    }
    return s && e;
}

/* Rust f64::min / f64::max: a NaN operand yields the other operand. */
static double rmin(double a, double b) { return isnan(a) ? b : isnan(b) ? a : (b < a ? b : a); }
static double rmax(double a, double b) { return isnan(a) ? b : isnan(b) ? a : (b > a ? b : a); }

/* ------------------------------------------------------------------ */
/* Chunk-level readers                                                 */
/* ------------------------------------------------------------------ */
/* precision >= 0: Decoder::with_precision (only BUFF decodes differently,
 * chunk_reader.rs:73-77). */
static int chunk_decompress_p(int scheme, const uint8_t *p, size_t size, uint64_t nrec, int32_t precision,
                              dvec *o);
static int chunk_decompress(int scheme, const uint8_t *p, size_t size, uint64_t nrec, dvec *o) {
    return chunk_decompress_p(scheme, p, size, nrec, -1, o);
}
static int chunk_decompress_p(int scheme, const uint8_t *p, size_t size, uint64_t nrec, int32_t precision,
                              dvec *o) {
    msb_t r = {p, size, 0, 0};
    switch (scheme) {
    case 0: /* chunk_reader/uncompressed.rs:56-71 (read_exact of n records) */
        if (nrec * 8 > size) return ORC_UNEXPECTED_END;
        for (uint64_t i = 0; i < nrec; ++i)
            if (dpush(o, bits2d(rd64(p + 8 * i)))) return ORC_NOMEM;
        return ORC_OK;
    case 1: { /* chunk_reader/run_length.rs:68-91 */
        if (size < 8) return ORC_UNEXPECTED_END;
        uint64_t fields = rd64(p);
        if (fields > (size - 8) / 9) return ORC_UNEXPECTED_END;
        for (uint64_t f = 0; f < fields; ++f) {
            uint8_t run = p[8 + 9 * f];
            double v = bits2d(rd64(p + 8 + 9 * f + 1));
            for (unsigned k = 0; k < run; ++k)
                if (dpush(o, v)) return ORC_NOMEM;
        }
        return ORC_OK;
    }
    case 2: return precision >= 0 ? dec_buff_precision(p, size, (uint32_t)precision, o) : dec_buff(p, size, o);
    case 3: return dec_gorilla(&r, o);
    case 4: return dec_chimp(&r, o);
    case 5: return dec_chimp128(&r, o);
    case 6: return dec_elf(&r, o, 1);
    case 7: return dec_elf(&r, o, 0);
    case 8: {
        spz_hdr h;
        int64_t *q = malloc((nrec ? nrec : 1) * sizeof(int64_t));
        if (!q) return ORC_NOMEM;
        int rc = spz_walk(&r, nrec, &h, q);
        for (uint64_t i = 0; !rc && i < nrec; ++i)
            if (dpush(o, (double)q[i] / (double)h.scale)) rc = ORC_NOMEM;
        free(q);
        return rc;
    }
    default:
        return ORC_UNSUPPORTED;
    }
}

static int bit_get(const uint64_t *bs, uint64_t i, uint64_t n) {
    return i < n && ((bs[i >> 6] >> (i & 63)) & 1);
}

typedef struct {
    orc_file f;
    const uint8_t *b;
    size_t len;
} ctx_t;

static int ctx_open(ctx_t *c, const uint8_t *b, size_t len) {
    c->b = b;
    c->len = len;
    return file_parse(b, len, &c->f);
}

/* Chunk selection (chunk_id) and pruning (bitmask ∩ range = ∅),
 * api/decoder.rs:216-226.  Builds the chunk-restricted mask if a bitmask is
 * given.  Returns 1 if the chunk is to be processed. */
static int chunk_selected(ctx_t *c, size_t k, const uint64_t *bs, int64_t chunk_id, uint8_t **mask) {
    *mask = NULL;
    if (chunk_id >= 0 && (uint64_t)chunk_id != k) return 0;
    if (!bs) return 1;
    uint64_t lo = c->f.logi[k], n = c->f.nrec[k];
    uint8_t *m = malloc(n ? n : 1);
    int any = 0;
    for (uint64_t i = 0; i < n; ++i) {
        m[i] = (uint8_t)bit_get(bs, lo + i, c->f.n_records);
        any |= m[i];
    }
    if (!any) { free(m); return 0; }
    *mask = m;
    return 1;
}

/* default_* of decoder/query.rs:15-154 iterate over the decoded values with
 * global index logical_offset + i and test bitmask.contains; decoded counts
 * beyond the chunk's footer count are tested against the global bitmask. */
static int in_mask(const uint64_t *bs, uint64_t n_records, uint64_t gid) {
    return !bs || bit_get(bs, gid, n_records);
}

int orc_file_records(const uint8_t *file, size_t len, uint64_t *n_records, uint64_t *n_chunks) {
    orc_file f;
    int rc = file_parse(file, len, &f);
    if (rc) return rc;
    if (n_records) *n_records = f.n_records;
    if (n_chunks) *n_chunks = f.n_chunks;
    file_free(&f);
    return ORC_OK;
}

int orc_materialize(const uint8_t *file, size_t len, const uint64_t *bs, int64_t chunk_id,
                    double *out, size_t cap, size_t *n_out) {
    return orc_materialize_p(file, len, bs, chunk_id, -1, out, cap, n_out);
}
int orc_materialize_p(const uint8_t *file, size_t len, const uint64_t *bs, int64_t chunk_id, int32_t precision,
                      double *out, size_t cap, size_t *n_out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    size_t no = 0;
    for (size_t k = 0; k < c.f.n_chunks && !rc; ++k) {
        uint8_t *m;
        if (!chunk_selected(&c, k, bs, chunk_id, &m)) continue;
        free(m);
        dvec d = {0};                   /* sized for the chunk up front, as Vec::with_capacity */
        d.cap = c.f.nrec[k] ? c.f.nrec[k] : 1;
        if (!(d.v = malloc(d.cap * sizeof(double)))) { rc = ORC_NOMEM; break; }
        rc = chunk_decompress_p(c.f.scheme, file + c.f.phys[k], c.f.size[k], c.f.nrec[k], precision, &d);
        for (size_t i = 0; !rc && i < d.n; ++i) {
            if (!in_mask(bs, c.f.n_records, c.f.logi[k] + i)) continue;
            if (no >= cap) { rc = ORC_CAPACITY; break; }
            out[no++] = d.v[i];
        }
        free(d.v);
    }
    *n_out = no;
    file_free(&c.f);
    return rc;
}

/* Per-chunk filter into dense flags res[0..nrec). */
static int chunk_filter(ctx_t *c, size_t k, const orc_pred *p, const uint64_t *bs,
                        const uint8_t *mask, uint8_t *res) {
    uint64_t n = c->f.nrec[k];
    const uint8_t *cp = c->b + c->f.phys[k];
    memset(res, 0, n);
    if (c->f.scheme == 2) {                       /* BUFFReader::filter */
        buff_t b;
        int rc = buff_open(cp, c->f.size[k], &b);
        if (rc) return rc;
        if (b.n != n) return ORC_CORRUPT;
        return buff_filter(&b, p, mask, res);
    }
    if (c->f.scheme == 8) {                       /* DeltaSprintzReader::filter :91-162 */
        msb_t r = {cp, c->f.size[k], 0, 0};
        spz_hdr h;
        int64_t *q = malloc((n ? n : 1) * sizeof(int64_t));
        int rc = spz_walk(&r, n, &h, q);
        if (!rc)
            for (uint64_t i = 0; i < n; ++i)
                res[i] = pred_eval_q(p, (double)h.scale, q[i]) && (!mask || mask[i]);
        free(q);
        return rc;
    }
    dvec d = {0};                                 /* default_filter :37-62 */
    int rc = chunk_decompress(c->f.scheme, cp, c->f.size[k], n, &d);
    for (size_t i = 0; !rc && i < d.n; ++i) {
        if (i >= n) { rc = ORC_CORRUPT; break; }
        if (!in_mask(bs, c->f.n_records, c->f.logi[k] + i)) continue;
        res[i] = (uint8_t)pred_eval_f(p, d.v[i]);
    }
    free(d.v);
    return rc;
}

int orc_filter(const uint8_t *file, size_t len, const orc_pred *p, const uint64_t *bs,
               int64_t chunk_id, uint64_t *out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    memset(out, 0, ((c.f.n_records + 63) / 64) * 8);
    for (size_t k = 0; k < c.f.n_chunks && !rc; ++k) {
        uint8_t *m;
        if (!chunk_selected(&c, k, bs, chunk_id, &m)) continue;
        uint64_t n = c.f.nrec[k];
        uint8_t *res = malloc(n ? n : 1);
        rc = chunk_filter(&c, k, p, bs, m, res);
        for (uint64_t i = 0; !rc && i < n; ++i)
            if (res[i]) {
                uint64_t g = c.f.logi[k] + i;
                out[g >> 6] |= 1ull << (g & 63);
            }
        free(res);
        free(m);
    }
    file_free(&c.f);
    return rc;
}

int orc_filter_materialize(const uint8_t *file, size_t len, const orc_pred *p, const uint64_t *bs,
                           int64_t chunk_id, double *out, size_t cap, size_t *n_out) {
    return orc_filter_materialize_p(file, len, p, bs, chunk_id, -1, out, cap, n_out);
}
int orc_filter_materialize_p(const uint8_t *file, size_t len, const orc_pred *p, const uint64_t *bs,
                             int64_t chunk_id, int32_t precision, double *out, size_t cap, size_t *n_out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    size_t no = 0;
    for (size_t k = 0; k < c.f.n_chunks && !rc; ++k) {
        uint8_t *m;
        if (!chunk_selected(&c, k, bs, chunk_id, &m)) continue;
        uint64_t n = c.f.nrec[k];
        dvec d = {0};
        rc = chunk_decompress_p(c.f.scheme, file + c.f.phys[k], c.f.size[k], n, precision, &d);
        uint8_t *res = NULL;
        if (!rc && c.f.scheme == 2) {
            /* BUFFReader::filter_materialize :230-241: in-situ filter, then
             * default_materialize with the filtered bitmask. */
            res = malloc(n ? n : 1);
            rc = chunk_filter(&c, k, p, bs, m, res);
        }
        for (size_t i = 0; !rc && i < d.n; ++i) {
            int keep;
            if (res) keep = i < n && res[i];
            else keep = in_mask(bs, c.f.n_records, c.f.logi[k] + i) && pred_eval_f(p, d.v[i]);
            if (!keep) continue;
            if (no >= cap) { rc = ORC_CAPACITY; break; }
            out[no++] = d.v[i];
        }
        free(res);
        free(d.v);
        free(m);
    }
    *n_out = no;
    file_free(&c.f);
    return rc;
}

/* op: 0 sum, 1 min, 2 max.  Per-chunk semantics of the codec readers, then
 * the Decoder's fold (api/decoder.rs:323-439). */
static int chunk_agg(ctx_t *c, size_t k, const uint64_t *bs, const uint8_t *mask, int op, double *res) {
    uint64_t n = c->f.nrec[k];
    const uint8_t *cp = c->b + c->f.phys[k];
    if (c->f.scheme == 2) {
        buff_t b;
        int rc = buff_open(cp, c->f.size[k], &b);
        if (rc) return rc;
        if (b.n != n) return ORC_CORRUPT;
        if (op == 0) { *res = buff_sum(&b, mask); return ORC_OK; }
        *res = buff_minmax(&b, mask, op == 1);
        return ORC_OK;
    }
    if (c->f.scheme == 8) {                 /* sprintz.rs:176-446 */
        msb_t r = {cp, c->f.size[k], 0, 0};
        spz_hdr h;
        int64_t *q = malloc((n ? n : 1) * sizeof(int64_t));
        int rc = spz_walk(&r, n, &h, q);
        if (!rc) {
            int64_t acc = op == 0 ? 0 : op == 1 ? INT64_MAX : INT64_MIN;
            uint64_t cnt = 0;
            for (uint64_t i = 0; i < n; ++i) {
                if (mask && !mask[i]) continue;
                ++cnt;
                if (op == 0) acc = (int64_t)((uint64_t)acc + (uint64_t)q[i]);
                else if (op == 1) acc = q[i] < acc ? q[i] : acc;
                else acc = q[i] > acc ? q[i] : acc;
            }
            *res = (mask && cnt == 0) ? NAN : (double)acc / (double)h.scale;
        }
        free(q);
        return rc;
    }
    dvec d = {0};
    int rc = chunk_decompress(c->f.scheme, cp, c->f.size[k], n, &d);
    double acc = op == 0 ? 0.0 : op == 1 ? INFINITY : -INFINITY;
    for (size_t i = 0; !rc && i < d.n; ++i) {
        if (!in_mask(bs, c->f.n_records, c->f.logi[k] + i)) continue;
        if (op == 0) acc += d.v[i];                   /* default_sum :91-110 */
        else if (op == 1) acc = rmin(acc, d.v[i]);    /* default_min :113-132 */
        else acc = rmax(acc, d.v[i]);                 /* default_max :135-154 */
    }
    free(d.v);
    *res = acc;
    return rc;
}

static int orc_agg(const uint8_t *file, size_t len, const uint64_t *bs, int64_t chunk_id, int op,
                   double *out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    double result = op == 0 ? 0.0 : op == 1 ? INFINITY : -INFINITY;
    for (size_t k = 0; k < c.f.n_chunks && !rc; ++k) {
        uint8_t *m;
        if (!chunk_selected(&c, k, bs, chunk_id, &m)) continue;
        double r = 0;
        rc = chunk_agg(&c, k, bs, m, op, &r);
        free(m);
        if (rc) break;
        if (op == 0) result += r;
        else if (op == 1) result = rmin(result, r);
        else result = rmax(result, r);
    }
    *out = result;
    file_free(&c.f);
    return rc;
}

int orc_sum(const uint8_t *f, size_t l, const uint64_t *bs, int64_t id, double *o) { return orc_agg(f, l, bs, id, 0, o); }
int orc_min(const uint8_t *f, size_t l, const uint64_t *bs, int64_t id, double *o) { return orc_agg(f, l, bs, id, 1, o); }
int orc_max(const uint8_t *f, size_t l, const uint64_t *bs, int64_t id, double *o) { return orc_agg(f, l, bs, id, 2, o); }

int orc_sum_compensated(const uint8_t *file, size_t len, const uint64_t *bs, int64_t chunk_id,
                        double *out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    double s = 0.0, comp = 0.0; /* Neumaier */
    for (size_t k = 0; k < c.f.n_chunks && !rc; ++k) {
        uint8_t *m;
        if (!chunk_selected(&c, k, bs, chunk_id, &m)) continue;
        free(m);
        dvec d = {0};
        rc = chunk_decompress(c.f.scheme, file + c.f.phys[k], c.f.size[k], c.f.nrec[k], &d);
        for (size_t i = 0; !rc && i < d.n; ++i) {
            if (!in_mask(bs, c.f.n_records, c.f.logi[k] + i)) continue;
            double x = d.v[i], t = s + x;
            if (fabs(s) >= fabs(x)) comp += (s - t) + x;
            else comp += (x - t) + s;
            s = t;
        }
        free(d.v);
    }
    *out = s + comp;
    file_free(&c.f);
    return rc;
}

/* ------------------------------------------------------------------ */
/* ML queries and multi-column aggregations                            */
/* ------------------------------------------------------------------ */

/* One decoded chunk (the reference's GeneralChunkReader for chunk k). */
static int chunk_values(ctx_t *c, size_t k, int32_t precision, dvec *d) {
    memset(d, 0, sizeof *d);
    return chunk_decompress_p(c->f.scheme, c->b + c->f.phys[k], c->f.size[k], c->f.nrec[k], precision, d);
}

/* default_distance_squared / default_dot_product, decoder/query.rs:157-192:
 * decompressed[min(len, offset)..] zipped with target, folded from -0.0 as
 * core's f64 Sum does. */
static double fold_zip(const dvec *d, size_t offset, const double *t, size_t tn, int dot) {
    size_t o = offset < d->n ? offset : d->n;
    size_t cnt = d->n - o < tn ? d->n - o : tn;
    double s = -0.0;
    for (size_t i = 0; i < cnt; ++i) {
        double a = d->v[o + i], b = t[i];
        if (dot) s += a * b;
        else { double x = a - b; s += x * x; }
    }
    return s;
}

int orc_chunk_fold(const uint8_t *file, size_t len, uint64_t chunk_id, size_t offset, const double *target,
                   size_t tn, int dot, double *out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    if (chunk_id >= c.f.n_chunks) { file_free(&c.f); return ORC_CORRUPT; }
    dvec d;
    rc = chunk_values(&c, chunk_id, -1, &d);
    if (!rc) *out = fold_zip(&d, offset, target, tn, dot);
    free(d.v);
    file_free(&c.f);
    return rc;
}

/* Decoder::knn1, api/decoder.rs:469-533, statement by statement (including
 * `offset += remaining_records` after remaining_records was reset, :503-505). */
int orc_knn1(const uint8_t *file, size_t len, const double *target, size_t tlen, uint64_t *index, double *dist) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    if (tlen == 0) { file_free(&c.f); return ORC_CORRUPT; } /* the loop would never end */
    uint64_t pidx = 0;
    double pdist = 0.0;
    uint64_t midx = UINT64_MAX;
    double mdist = INFINITY;
    size_t remaining = 0;
    for (size_t k = 0; k < c.f.n_chunks && !rc; ++k) {
        size_t offset = 0, n = (size_t)c.f.nrec[k];
        dvec d;
        rc = chunk_values(&c, k, -1, &d);
        if (rc) { free(d.v); break; }
        if (remaining != 0) {
            size_t start = tlen - remaining;
            pdist += fold_zip(&d, 0, target + start, tlen - start, 0);
            if (pdist < mdist) { midx = pidx; mdist = pdist; }
            remaining = 0;
            pidx += 1;
            offset += remaining;
        }
        while (offset < n) {
            size_t nin = tlen < n - offset ? tlen : n - offset;
            pdist = fold_zip(&d, offset, target, nin, 0);
            if (nin == tlen) {
                if (pdist <= mdist) { midx = pidx; mdist = pdist; }
                pidx += 1;
            } else {
                remaining = tlen - nin;
            }
            offset += nin;
        }
        free(d.v);
    }
    *index = midx;
    *dist = mdist;
    file_free(&c.f);
    return rc;
}

/* Decoder::matmul, api/decoder.rs:606-656 (target column-first). */
int orc_matmul_dot(const uint8_t *file, size_t len, const double *target, size_t trows, size_t tcols,
                   size_t drows, size_t dcols, double *out) {
    ctx_t c;
    int rc = ctx_open(&c, file, len);
    if (rc) return rc;
    if (!drows || !dcols) { file_free(&c.f); return ORC_CORRUPT; }
    if (dcols != trows || c.f.n_records % (drows * dcols) != 0) { file_free(&c.f); return 3; }
    uint64_t batch = c.f.n_records / (drows * dcols);
    size_t off = 0, ci = 0, no = 0;
    dvec d;
    if (c.f.n_chunks == 0) { file_free(&c.f); return batch ? 3 : ORC_OK; }
    rc = chunk_values(&c, 0, -1, &d);
    for (uint64_t b = 0; b < batch && !rc; ++b) {
        if (off >= c.f.nrec[ci]) {
            off = 0;
            ci += 1;
            free(d.v);
            if (ci >= c.f.n_chunks) { d.v = NULL; rc = 3; break; }
            rc = chunk_values(&c, ci, -1, &d);
            if (rc) break;
        }
        for (size_t r = 0; r < drows; ++r)
            for (size_t j = 0; j < tcols; ++j)
                out[no++] = fold_zip(&d, off + r * dcols, target + j * trows, trows, 1);
        off += drows * dcols;
    }
    free(d.v);
    file_free(&c.f);
    return rc;
}

/* Expr::compute_chunk (decoder/expr.rs:67-123) over postfix nodes (kind 0
 * column, 1 literal, 2 add, 3 sub, 4 mul, 5 div), recursively as the
 * reference: a column is its chunk's records selected by the bitmask, a
 * binary node zips its operands (or maps a literal over one). */
typedef struct { int32_t kind; uint32_t index; double literal; } orc_node;
typedef struct { double *v; size_t n; int lit; double litv; } orc_val;

static int expr_eval(const orc_node *nodes, size_t n_nodes, dvec *chunk_sel, orc_val *res) {
    orc_val st[64];
    size_t sp = 0;
    for (size_t p = 0; p < n_nodes; ++p) {
        const orc_node *e = &nodes[p];
        if (e->kind == 0) {
            orc_val v = {malloc((chunk_sel[e->index].n ? chunk_sel[e->index].n : 1) * sizeof(double)),
                         chunk_sel[e->index].n, 0, 0.0};
            memcpy(v.v, chunk_sel[e->index].v, v.n * sizeof(double));
            st[sp++] = v;
        } else if (e->kind == 1) {
            orc_val v = {NULL, 0, 1, e->literal};
            st[sp++] = v;
        } else {
            orc_val b = st[--sp], a = st[--sp];
            if (a.lit && b.lit) return 3; /* PreconditionsNotMet */
            size_t n = a.lit ? b.n : b.lit ? a.n : (a.n < b.n ? a.n : b.n);
            orc_val r = {malloc((n ? n : 1) * sizeof(double)), n, 0, 0.0};
            for (size_t i = 0; i < n; ++i) {
                double x = a.lit ? a.litv : a.v[i], y = b.lit ? b.litv : b.v[i];
                r.v[i] = e->kind == 2 ? x + y : e->kind == 3 ? x - y : e->kind == 4 ? x * y : x / y;
            }
            free(a.v);
            free(b.v);
            st[sp++] = r;
        }
    }
    *res = st[0];
    return res->lit ? 3 : ORC_OK;
}

/* Aggregation::compute_multiple (api/decoder.rs:874-917) of one
 * aggregation over files[0..n_files): kind 0 none (values), 1 max, 2 min,
 * 3 sum.  A bare column takes the decoder's own sum / min / max.  sum_comp:
 * a compensated (Neumaier) sum of the same records. */
int orc_aggregate(const uint8_t *const *files, const size_t *lens, size_t n_files, const void *nodes_v,
                  size_t n_nodes, int kind, const uint64_t *bs, double *out_scalar, double *sum_comp,
                  double *out_values, size_t cap, size_t *n_out) {
    const orc_node *nodes = (const orc_node *)nodes_v;
    if (n_nodes == 1 && nodes[0].kind == 0) {
        if (kind == 0) return ORC_UNSUPPORTED;
        int op = kind == 3 ? 0 : kind == 2 ? 1 : 2;
        int rc = orc_agg(files[nodes[0].index], lens[nodes[0].index], bs, -1, op, out_scalar);
        if (!rc && sum_comp) rc = orc_sum_compensated(files[nodes[0].index], lens[nodes[0].index], bs, -1, sum_comp);
        return rc;
    }
    ctx_t *cs = calloc(n_files, sizeof(ctx_t));
    int rc = ORC_OK;
    for (size_t i = 0; i < n_files && !rc; ++i) rc = ctx_open(&cs[i], files[i], lens[i]);
    size_t nc = rc ? 0 : (size_t)cs[0].f.n_chunks;
    double result = kind == 1 ? -INFINITY : kind == 2 ? INFINITY : 0.0;
    double s = 0.0, comp = 0.0;
    size_t no = 0;
    for (size_t k = 0; k < nc && !rc; ++k) {
        dvec *sel = calloc(n_files, sizeof(dvec));
        for (size_t i = 0; i < n_files && !rc; ++i) {
            if (k >= cs[i].f.n_chunks) { rc = 3; break; }
            dvec d;
            rc = chunk_values(&cs[i], k, -1, &d);
            for (size_t j = 0; !rc && j < d.n; ++j) /* default_materialize :15-34 */
                if (in_mask(bs, cs[i].f.n_records, cs[i].f.logi[k] + j)) dpush(&sel[i], d.v[j]);
            free(d.v);
        }
        orc_val r = {0};
        if (!rc) rc = expr_eval(nodes, n_nodes, sel, &r);
        if (!rc) {
            if (kind == 0) {
                for (size_t j = 0; j < r.n && !rc; ++j) {
                    if (no >= cap) rc = ORC_CAPACITY;
                    else out_values[no++] = r.v[j];
                }
            } else if (kind == 3) {
                double cs_ = -0.0;  /* iter().sum::<f64>() */
                for (size_t j = 0; j < r.n; ++j) {
                    cs_ += r.v[j];
                    double x = r.v[j], t = s + x;
                    if (fabs(s) >= fabs(x)) comp += (s - t) + x;
                    else comp += (x - t) + s;
                    s = t;
                }
                result += cs_;
            } else {
                double m = kind == 1 ? -INFINITY : INFINITY;
                for (size_t j = 0; j < r.n; ++j) m = kind == 1 ? rmax(m, r.v[j]) : rmin(m, r.v[j]);
                result = kind == 1 ? rmax(result, m) : rmin(result, m);
            }
        }
        free(r.v);
        for (size_t i = 0; i < n_files; ++i) free(sel[i].v);
        free(sel);
    }
    if (kind == 0) *n_out = no;
    else *out_scalar = result;
    if (sum_comp) *sum_comp = s + comp;
    for (size_t i = 0; i < n_files; ++i) file_free(&cs[i].f);
    free(cs);
    return rc;
}
