/*
 * swm_oracle.c -- CPU restatement of SWMapper's mapping hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Not part of the product library;
 * loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
 *
 * Each function restates one reference function, statement by statement where
 * the statement order changes results, and cites the reference file:line.
 * "Define" decisions for undefined reference behaviour:
 *   Q1  a u32 shift by >= 32 (or by a negative count) yields 0.
 *   Q2  genome words past the end of .base_2bits read as 0.
 *   Q5  long seeds beyond floor(L/15) contribute no hits.
 *   Q8  a batch whose blocks are all empty after the 15-mer filter yields 0
 *       candidates.
 *   Q12/Q13 candidate windows are the true genome bits at any length (the
 *       reference's fixed 16-word lane slots and stale tails are not modelled).
 *   Q23 (index) slots of .positions / .pos_lseeds that the scatter never
 *       writes hold 0 (the reference leaves malloc memory there).
 *
 * ORC_MUTANT (0 unless set) builds a copy with one planted mistake, for the
 * tests that show a known answer fails a plausible wrong restatement:
 *   1 the textbook seed-heap sift-down (a lone left child compared)
 *   2 the seed heap keyed by read-start-adjusted positions
 *   3 a batch cut one chunk early (pos_size + 32 < 2048)
 *   4 an insertion at the traceback's first step labelled 'I'
 *   5 the traceback trying an insertion before a match
 */
#include "swm_oracle.h"

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

#include <ctype.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Q1: funnel-shift intent, a shift count outside [0, 31] gives 0. */
static inline uint32_t shl32(uint32_t x, int64_t s) { return (s < 0 || s >= 32) ? 0u : (x << s); }
static inline uint32_t shr32(uint32_t x, int64_t s) { return (s < 0 || s >= 32) ? 0u : (x >> s); }

/* ======================================================================== */
/* Index build: serial_index.c (HOST_BUILD, :26)                            */
/* ======================================================================== */

/* Per-base state of the serial builder (globals at serial_index.c:31-53). */
typedef struct {
    uint32_t position;         /* :31 global base counter */
    uint64_t k;                /* :32 seed_length */
    uint64_t cur_seed_length;  /* :33 */
    uint64_t seed;             /* :34 rolling code */
    uint32_t enc_lbits, enc_hbits, enc_seeds_2bits, enc_3bits; /* :49-53 */
    uint32_t *seed_count, *base_2bits, *seeds_2bits, *base_3bits;
    uint64_t valid_seed_size;  /* :37 */
} ix_state_t;

/* One base of get_seed()'s loop (serial_index.c:160-235) followed by the
 * validity/count step of get_seed()/chromo_hash_index() (:236-244, :261-264).
 * The contigs are processed back to back with the state carried over, so the
 * concatenated sequence yields the same sequence of state changes (Q17). */
static void ix_push_base(ix_state_t *st, char raw)
{
    const uint64_t k = st->k;
    const uint32_t p = st->position;
    int code;
    switch (toupper((unsigned char)raw)) {
    case 'A': code = 0; break;
    case 'C': code = 1; break;
    case 'G': code = 2; break;
    case 'T': code = 3; break;
    default:  code = -1; break;
    }
    if (code >= 0) {
        st->seed = (st->seed >> 2) | ((uint64_t)code << ((k - 1) * 2));   /* :166,176,186,195 */
        st->cur_seed_length++;
        st->enc_lbits |= (uint32_t)(code & 1) << (p % 32);                /* :169-199 */
        st->enc_hbits |= (uint32_t)(code >> 1) << (p % 32);
        st->enc_seeds_2bits |= (uint32_t)code << ((p % 16) * 2);
    } else {                                                              /* :203-214 */
        st->enc_3bits |= 1u << (p % 32);
        st->cur_seed_length = 0;
        st->seed = 0;
    }
    if ((p + 1) % 32 == 0) {                                              /* :217-222 */
        st->base_2bits[2 * (p / 32)] = st->enc_lbits;
        st->base_2bits[2 * (p / 32) + 1] = st->enc_hbits;
        st->enc_lbits = 0;
        st->enc_hbits = 0;
    }
    if ((p + 1) % 16 == 0) {                                              /* :223-228 */
        st->seeds_2bits[p / 16] = st->enc_seeds_2bits;
        st->enc_seeds_2bits = 0;
    }
    if ((p + 1) % 32 == 0) {                                              /* :229-232 */
        st->base_3bits[p / 32] = st->enc_3bits;
        st->enc_3bits = 0;
    }
    st->position++;                                                       /* :234 */
    if (st->cur_seed_length == k) {                                       /* :236-238 */
        st->cur_seed_length--;
        /* a complete seed was returned; count it when its start is on the
         * SEED_STEP grid (chromo_hash_index, :261-264) */
        if (((uint64_t)st->position - k) % ORC_SEED_STEP == 0) {
            st->seed_count[st->seed]++;
            st->valid_seed_size++;
        }
    }
}

/* get_seed_from_enc_bits (serial_index.c:280-331).  Returns the seed code
 * or 0xffffffff, and the 12-bit residual code through *lseed. */
static uint32_t ix_seed_from_enc_bits(const ix_state_t *st, uint64_t pos, uint16_t *lseed)
{
    const uint64_t k = st->k;
    uint32_t seed_mask = (uint32_t)((1ull << (k * 2)) - 1);
    uint32_t flag_mask = (uint32_t)((1ull << k) - 1);
    uint32_t v = (shr32(st->seeds_2bits[pos / 16], (pos % 16) * 2) |
                  shl32(st->seeds_2bits[pos / 16 + 1], (16 - pos % 16) * 2)) & seed_mask; /* :286 */
    const uint32_t res_len = ORC_SEED_STEP - 1;
    const uint32_t mask = (1u << res_len) - 1;
    uint64_t s_pos = 0;
    if (pos >= res_len) s_pos = pos - res_len;                             /* :293-294, Q16 */
    uint32_t shift = s_pos % 32;
    s_pos /= 32;
    uint32_t fl = shr32(st->base_2bits[s_pos * 2], shift) | shl32(st->base_2bits[(s_pos + 1) * 2], 32 - (int64_t)shift);
    uint32_t fh = shr32(st->base_2bits[s_pos * 2 + 1], shift) | shl32(st->base_2bits[(s_pos + 1) * 2 + 1], 32 - (int64_t)shift);
    fl &= mask;
    fh &= mask;
    s_pos = pos + k;                                                       /* :303 */
    shift = s_pos % 32;
    s_pos /= 32;
    uint32_t bl = shr32(st->base_2bits[s_pos * 2], shift) | shl32(st->base_2bits[(s_pos + 1) * 2], 32 - (int64_t)shift);
    uint32_t bh = shr32(st->base_2bits[s_pos * 2 + 1], shift) | shl32(st->base_2bits[(s_pos + 1) * 2 + 1], 32 - (int64_t)shift);
    bl &= mask;
    bh &= mask;
    fl |= bl << res_len;                                                   /* :312-315 */
    fh |= bh << res_len;
    *lseed = (uint16_t)(fl | (fh << (res_len * 2)));
    uint32_t flag = (shr32(st->base_3bits[pos / 32], pos % 32) |
                     shl32(st->base_3bits[pos / 32 + 1], 32 - (int64_t)(pos % 32))) & flag_mask; /* :321 */
    if (flag != 0) v = 0xffffffffu;
    return v;
}

int orc_index_build(const char *seq, uint64_t n, uint32_t k, orc_index_t *out)
{
    memset(out, 0, sizeof(*out));
    if (!seq || k < 2 || k > 15 || n < k || n >= 0xffffffffull) return -1;
    const uint64_t nbuckets = 1ull << (2 * k);
    const uint64_t nwords = (n + 31) / 32;
    ix_state_t st;
    memset(&st, 0, sizeof(st));
    st.k = k;
    /* construct_hash_index (serial_index.c:359-445).  base_2bits and
     * seeds_2bits are zeroed by the reference (:377-378); base_3bits is not,
     * and words past the written ones read as 0 here (Q2). */
    st.seed_count = (uint32_t *)calloc(nbuckets, sizeof(uint32_t));
    st.base_2bits = (uint32_t *)calloc(2 * nwords + 64, sizeof(uint32_t));
    st.seeds_2bits = (uint32_t *)calloc(2 * nwords + 64, sizeof(uint32_t));
    st.base_3bits = (uint32_t *)calloc(nwords + 64, sizeof(uint32_t));
    uint32_t *hash_index = (uint32_t *)calloc(nbuckets + 1, sizeof(uint32_t));
    if (!st.seed_count || !st.base_2bits || !st.seeds_2bits || !st.base_3bits || !hash_index) goto fail;

    for (uint64_t q = 0; q < n; q++) ix_push_base(&st, seq[q]);            /* chromo_hash_index per contig */

    /* final flush of the partial words (:423-426).  When position is a
     * multiple of 32 (16 for seeds_2bits) the accumulator is empty and this
     * overwrites the last complete word with 0 -- reproduced as written. */
    {
        uint32_t pos = st.position;
        st.base_2bits[2 * ((pos + 31) / 32 - 1)] = st.enc_lbits;
        st.base_2bits[2 * ((pos + 31) / 32 - 1) + 1] = st.enc_hbits;
        st.base_3bits[(pos + 31) / 32 - 1] = st.enc_3bits;
        st.seeds_2bits[(pos + 15) / 16 - 1] = st.enc_seeds_2bits;
    }

    /* store_positions (:332-356): exclusive scan then stable scatter. */
    uint64_t M = st.valid_seed_size;
    out->positions = (uint32_t *)calloc(M + 64, sizeof(uint32_t));        /* Q23: unwritten slots = 0 */
    out->pos_lseeds = (uint16_t *)calloc(M + 64, sizeof(uint16_t));
    if (!out->positions || !out->pos_lseeds) goto fail;
    hash_index[0] = 0;
    for (uint64_t i = 0; i < nbuckets; i++) hash_index[i + 1] = hash_index[i] + st.seed_count[i];
    for (uint64_t i = 0; i < (uint64_t)st.position - k + 1; i += ORC_SEED_STEP) {
        uint16_t ls;
        uint32_t v = ix_seed_from_enc_bits(&st, i, &ls);
        if (v != 0xffffffffu) {
            uint64_t slot = hash_index[v];
            if (slot < M + 64) {   /* the reference's +8 pad absorbs the overflow */
                out->positions[slot] = (uint32_t)i;
                out->pos_lseeds[slot] = ls;
            }
            hash_index[v]++;
        }
    }

    out->k = k;
    out->n_bases = st.position;
    out->n_words = nwords;
    out->base_2bits = st.base_2bits;
    out->base_3bits = st.base_3bits;
    out->hash_count = st.seed_count;
    out->hash_offset_end = hash_index;
    out->n_pos = M;
    free(st.seeds_2bits);
    return 0;
fail:
    free(st.seed_count); free(st.base_2bits); free(st.seeds_2bits); free(st.base_3bits);
    free(hash_index); free(out->positions); free(out->pos_lseeds);
    memset(out, 0, sizeof(*out));
    return -2;
}

void orc_index_free(orc_index_t *ix)
{
    if (!ix) return;
    free(ix->base_2bits); free(ix->base_3bits); free(ix->hash_count);
    free(ix->hash_offset_end); free(ix->positions); free(ix->pos_lseeds);
    memset(ix, 0, sizeof(*ix));
}

/* ======================================================================== */
/* Mapper: mapper_cpe.c                                                     */
/* ======================================================================== */

/* get_seed_from_reads (mapper_cpe.c:1246-1364).  w[0..15] interleaved 2-bit
 * words, w[16..31] bit planes (low, high) per 32 bases. */
static uint32_t rd_get_seed(const char *rs, uint32_t *w, uint32_t *enc, uint32_t pre_seed,
                            uint64_t k, uint32_t seed_n, uint32_t L)
{
    uint64_t i = k - 1;
    if (seed_n % ORC_SEED_STEP == 0) i = 0;
    uint64_t roff = (uint64_t)(seed_n / ORC_SEED_STEP) * (k + ORC_SEED_STEP - 1);
    uint64_t soff = seed_n % ORC_SEED_STEP;
    uint32_t seed = pre_seed;
    uint64_t rp = roff + soff + i;
    for (; i < k; i++) {
        char b = rs[roff + soff + i];
        rp = roff + soff + i;
        uint32_t cs, cp;   /* seed code, plane/interleaved code */
        switch (b) {
        case 'A': cs = 0; cp = 0; break;
        case 'C': cs = 1; cp = 1; break;
        case 'G': cs = 2; cp = 2; break;
        case 'T': cs = 3; cp = 3; break;
        default:  cs = (uint32_t)(i % 4); cp = 1; break;              /* :1293-1300, Q15 */
        }
        seed = (uint32_t)(((uint64_t)seed >> 2) | ((uint64_t)cs << ((k - 1) * 2)));
        *enc |= cp << ((rp % 16) * 2);
        w[16 + (rp >> 5) * 2] |= (cp & 1u) << (rp & 0x1f);
        w[16 + (rp >> 5) * 2 + 1] |= (cp >> 1) << (rp & 0x1f);
        if (rp % 16 == 15) {                                               /* :1302-1305 */
            w[rp / 16] = *enc;
            *enc = 0;
        }
    }
    if ((uint64_t)L - rp < k) {                                            /* :1312 tail branch (Q3) */
        for (; rp < L; rp++) {
            char b = rs[rp];
            uint32_t cp;
            switch (b) {
            case 'A': cp = 0; break;
            case 'C': cp = 1; break;
            case 'G': cp = 2; break;
            case 'T': cp = 3; break;
            default:  cp = 1; break;
            }
            *enc |= cp << ((rp % 16) * 2);
            w[16 + (rp >> 5) * 2] |= (cp & 1u) << (rp & 0x1f);
            w[16 + (rp >> 5) * 2 + 1] |= (cp >> 1) << (rp & 0x1f);
            if (rp % 16 == 15) {
                w[rp / 16] = *enc;
                *enc = 0;
            }
        }
        w[(rp - 1) / 16] = *enc;                                           /* :1361 (enc not reset) */
    }
    return seed;
}

typedef struct { uint32_t count; uint32_t index; } hc_t;   /* hash_count_t, mapper_cpe.c:78-81 */

/* Per-strand working set (get_positions locals, mapper_cpe.c:2503-2515). */
typedef struct {
    uint32_t off2[128];      /* hash_offset[seed], hash_offset[seed+1] per seed */
    uint32_t cnt[64];        /* hash_count per seed */
    uint32_t w[32];          /* reads_2bits */
    uint32_t seeds[64];
    uint8_t  res[128];       /* res_seeds */
    hc_t     info[16];       /* hash_seed_info */
    uint32_t n_seeds;        /* seeds_size */
} strand_ws_t;

/* choose_seed (mapper_cpe.c:1365-1401). */
static void rd_choose_seed(const orc_ref_t *ref, const char *rs, uint32_t L, strand_ws_t *ws)
{
    const uint64_t k = ref->k;
    uint64_t seeds_size = (L / (k + ORC_SEED_STEP - 1)) * ORC_SEED_STEP;
    uint32_t seed = 0, enc = 0;
    for (uint64_t j = 0; j < seeds_size; j++) {
        seed = rd_get_seed(rs, ws->w, &enc, seed, k, (uint32_t)j, L);
        ws->seeds[j] = seed;
        ws->off2[j * 2] = ref->hash_offset[seed];                          /* :1375 */
        ws->off2[j * 2 + 1] = ref->hash_offset[seed + 1];
    }
    for (uint64_t j = 0; j < seeds_size; j++) ws->cnt[j] = ws->off2[j * 2 + 1] - ws->off2[j * 2];
    for (uint32_t j = 0; j < 16; j++) { ws->info[j].count = 0; ws->info[j].index = j; }
    for (uint64_t j = 0; j < seeds_size; j++) ws->info[j / ORC_SEED_STEP].count += ws->cnt[j];
    for (uint64_t j = 0; j < seeds_size / ORC_SEED_STEP; j++)            /* :1391-1399 exchange sort (Q4) */
        for (uint64_t q = j + 1; q < seeds_size / ORC_SEED_STEP; q++)
            if (ws->info[j].count > ws->info[q].count) {
                hc_t t = ws->info[j]; ws->info[j] = ws->info[q]; ws->info[q] = t;
            }
    ws->n_seeds = (uint32_t)seeds_size;
}

/* set_read_long_seeds (mapper_cpe.c:2451-2499): 3-base residuals. */
static void rd_long_seeds(uint32_t k, uint32_t L, strand_ws_t *ws)
{
    const uint32_t *w = ws->w;
    int seed_len = (int)k + ORC_SEED_STEP - 1;
    int seed_num = (int)L / seed_len;
    uint32_t seed_mask = (1u << seed_len) - 1;
    uint32_t res_mask = (1u << (ORC_SEED_STEP - 1)) - 1;
    for (int i = 0; i < seed_num; i++) {
        uint64_t seed_off = (uint64_t)i * seed_len;
        uint64_t j = seed_off >> 5;
        uint64_t sh = seed_off & 0x1f;
        /* j + 1 <= 7 for L <= 254, so w[16 + (j+1)*2 + 1] stays inside w[32] */
        uint32_t l_read = (shr32(w[16 + j * 2], sh) | shl32(j * 2 + 2 < 16 ? w[16 + (j + 1) * 2] : 0, 32 - (int64_t)sh)) & seed_mask;
        uint32_t h_read = (shr32(w[16 + j * 2 + 1], sh) | shl32(j * 2 + 3 < 16 ? w[16 + (j + 1) * 2 + 1] : 0, 32 - (int64_t)sh)) & seed_mask;
        for (uint32_t q = 0; q < ORC_SEED_STEP; q++) {
            uint32_t shift = k + q;
            uint32_t mask = (1u << q) - 1;
            uint32_t f = l_read & mask, b = (l_read >> shift) << q;
            ws->res[2 * (ORC_SEED_STEP * i + q)] = (uint8_t)((f | b) & res_mask);
            f = h_read & mask; b = (h_read >> shift) << q;
            ws->res[2 * (ORC_SEED_STEP * i + q) + 1] = (uint8_t)((f | b) & res_mask);
        }
    }
}

void orc_reverse_read(char *out, const char *read, uint32_t L)      /* mapper_cpe.c:2429-2450 */
{
    uint64_t j = (uint64_t)L - 1;
    for (uint64_t i = 0; i < L; i++) {
        char c;
        switch (read[i]) {
        case 'A': c = 'T'; break;
        case 'C': c = 'G'; break;
        case 'G': c = 'C'; break;
        case 'T': c = 'A'; break;
        default:  c = read[i]; break;
        }
        out[j - i] = c;
    }
}

int orc_encode_read(const char *read, uint32_t L, uint32_t k, uint32_t *words32, uint32_t *seeds)
{
    memset(words32, 0, 32 * sizeof(uint32_t));
    if (L > ORC_READ_LEN - 2 || L / (k + ORC_SEED_STEP - 1) > 32) return -1;
    uint64_t seeds_size = (L / (k + ORC_SEED_STEP - 1)) * ORC_SEED_STEP;
    uint32_t seed = 0, enc = 0;
    for (uint64_t j = 0; j < seeds_size; j++) {
        seed = rd_get_seed(read, words32, &enc, seed, k, (uint32_t)j, L);
        if (seeds) seeds[j] = seed;
    }
    return (int)seeds_size;
}

/* ---- candidate window (mapper_cpe.c:1403-1435, 1509-1586) ---- */

/* adjust_ref_len (mapper_cpe.c:1403-1423). */
static uint64_t adjust_ref_len(const orc_ref_t *ref, uint64_t e, uint64_t *shift, uint32_t base_pos, uint32_t L)
{
    uint64_t dist_error = e;
    uint64_t shift_tmp = e;
    uint64_t ref_len = L + 2 * shift_tmp;
    if (base_pos <= shift_tmp) {
        shift_tmp = base_pos;
        ref_len -= (dist_error - base_pos);
    }
    int64_t front = (int64_t)ref->ref_length - (int64_t)((uint64_t)base_pos - dist_error);
    int64_t back = (int64_t)((uint64_t)base_pos + L + dist_error) - (int64_t)ref->ref_length;
    if (front > back) {
        if (back >= 0) ref_len -= (uint64_t)(back + 1);
    } else {
        shift_tmp = dist_error - (uint64_t)front - 1;
        ref_len -= (uint64_t)front;
    }
    *shift = shift_tmp;
    return ref_len;
}

/* pos_start (mapper_cpe.c:1424-1435): window start or 0xffffffff. */
static uint32_t pos_start(const orc_ref_t *ref, uint64_t e, uint32_t base_pos, uint64_t *shift,
                          uint32_t *ref_len, uint32_t L)
{
    uint64_t t = adjust_ref_len(ref, e, shift, base_pos, L);
    *ref_len = (uint32_t)t;
    if ((uint64_t)*ref_len < (uint64_t)L - e) return 0xffffffffu;
    return (uint32_t)(base_pos - *shift);
}

/* Genome plane word starting at absolute bit position g (Q2, Q12, Q13). */
static uint32_t genome_word(const orc_ref_t *ref, int64_t g, int plane)
{
    if (g <= -32) return 0;
    uint64_t nw = ref->n_2bits_words / 2;
    if (g < 0) {
        uint32_t w0 = ref->base_2bits[plane];
        return shl32(w0, -g);
    }
    uint64_t wi = (uint64_t)g >> 5, sh = (uint64_t)g & 31;
    uint32_t a = wi < nw ? ref->base_2bits[2 * wi + plane] : 0;
    uint32_t b = wi + 1 < nw ? ref->base_2bits[2 * (wi + 1) + plane] : 0;
    return shr32(a, sh) | shl32(b, 32 - (int64_t)sh);
}

/* ---- banded Myers (semantics of mapper_cpe.c:519-660, compiled form
 * :668-976; the two agree for any word width >= 2e+2, Q14) ---- */
#define MYERS_BODY(WT)                                                               \
    int dist_error = (int)e;                                                         \
    WT band_length = 2 * (WT)e + 1;                                                  \
    WT band_down = 2 * (WT)e;                                                        \
    WT peq[4] = {0, 0, 0, 0};                                                        \
    WT tmp = 1;                                                                      \
    for (WT j = 0; j < band_length; j++) {                                           \
        uint32_t b1 = (win_lo[j >> 5] >> (j & 0x1f)) & 1u;                           \
        uint32_t b2 = (win_hi[j >> 5] >> (j & 0x1f)) & 1u;                           \
        peq[b1 | (b2 << 1)] |= tmp;                                                  \
        tmp <<= 1;                                                                   \
    }                                                                                \
    int last_high = (int)band_length - (int)L + (int)ref_len - (int)band_down - 1;   \
    WT vp = 0, vn = 0, x, d0, hn, hp;                                                \
    int err = 0;                                                                     \
    WT i_bd = band_down;                                                             \
    WT mask = (WT)1 << (band_length - 1);                                            \
    for (uint32_t ii = 0; ii * 16 < L; ii++) {                                       \
        uint32_t res = 16;                                                           \
        if ((ii + 1) * 16 > L) res = L - ii * 16;                                    \
        for (uint32_t i = 0; i < res; i++) {                                         \
            uint32_t be = (read_words[ii] >> (i * 2)) & 3u;                          \
            x = peq[be] | vn;                                                        \
            d0 = x & vp;                                                             \
            d0 = d0 + vp;                                                            \
            d0 = d0 ^ vp;                                                            \
            d0 = d0 | x;                                                             \
            hn = vp & d0;                                                            \
            hp = ~(vp | d0);                                                         \
            hp = vn | hp;                                                            \
            x = d0 >> 1;                                                             \
            vn = x & hp;                                                             \
            vp = ~(x | hp);                                                          \
            vp = hn | vp;                                                            \
            if (!(d0 & 1)) ++err;                                                    \
            peq[0] >>= 1; peq[1] >>= 1; peq[2] >>= 1; peq[3] >>= 1;                  \
            ++i_bd;                                                                  \
            uint32_t rw = (uint32_t)(i_bd >> 5);                                     \
            uint32_t e1 = (win_lo[rw] >> (i_bd & 0x1f)) & 1u;                        \
            uint32_t e2 = (win_hi[rw] >> (i_bd & 0x1f)) & 1u;                        \
            peq[e1 | (e2 << 1)] |= mask;                                             \
        }                                                                            \
    }                                                                                \
    int site = (int)ref_len - last_high - 1;                                         \
    int location = -1, error = ORC_MAXERROR;                                         \
    if (err <= dist_error && err < error) { error = err; location = site; }          \
    for (int i = 0; i < last_high; i++) {                                            \
        err = err + (int)((vp >> i) & (WT)1);                                        \
        err = err - (int)((vn >> i) & (WT)1);                                        \
        if (err <= dist_error && err < error) { error = err; location = site + i + 1; } \
    }                                                                                \
    *out_err = error;                                                                \
    *out_loc = location;

void orc_myers(const uint32_t *read_words, const uint32_t *win_lo, const uint32_t *win_hi,
               uint32_t L, uint32_t ref_len, uint32_t e, int32_t *out_err, int32_t *out_loc)
{
    MYERS_BODY(uint32_t)
}

void orc_myers64(const uint32_t *read_words, const uint32_t *win_lo, const uint32_t *win_hi,
                 uint32_t L, uint32_t ref_len, uint32_t e, int32_t *out_err, int32_t *out_loc)
{
    MYERS_BODY(uint64_t)
}

/* ---- record emission (mapper_cpe.c:214-517, 1437-1507) ---- */

typedef struct {
    /* cigar_t (mapper_cpe.c:98-108) */
    uint32_t cigar_len;
    uint32_t cigar_num;
    uint32_t location[ORC_MAX_CIGAR * 2 + 8];
    int main_flag;
    /* sinks */
    uint32_t read, strand;
    uint32_t block;
    uint32_t *read_count;      /* ref_info_s->cigar_size */
    orc_record_t *recs; uint64_t n_recs, cap_recs;
    int keep_records;
    /* per-read best tracking */
    orc_hit_t *hit;
} emit_t;

static int ndigits(uint32_t v) { int n = 0; while (v) { n++; v /= 10; } return n; }  /* sprint_length :214-228 */

/* sprint_length (mapper_cpe.c:214-228): the decimal digits of v (none for 0). */
static int sprint_length(char *out, uint32_t v)
{
    int n = ndigits(v), k = n;
    while (v) { out[--k] = (char)('0' + v % 10); v /= 10; }
    return n;
}

static void emit_record(emit_t *em, uint32_t loc, int dist, uint32_t text_len, const char *cigar)
{
    const int primary = em->main_flag && em->cigar_num == 0;   /* first record of the strand's first block */
    em->location[em->cigar_num++] = loc;
    em->cigar_len += text_len;
    if (em->hit) {
        orc_hit_t *h = em->hit;
        if (h->count == 0 || (uint32_t)dist < h->dist) {
            h->pos = loc; h->strand = (uint8_t)em->strand; h->dist = (uint8_t)dist;
        }
        h->count++;
    }
    if (em->keep_records) {
        if (em->n_recs == em->cap_recs) {
            em->cap_recs = em->cap_recs ? em->cap_recs * 2 : 64;
            em->recs = (orc_record_t *)realloc(em->recs, em->cap_recs * sizeof(orc_record_t));
        }
        orc_record_t *r = &em->recs[em->n_recs++];
        r->read = em->read; r->strand = (uint8_t)em->strand; r->dist = (uint8_t)dist;
        r->block = (uint16_t)em->block; r->loc = loc; r->text_len = text_len;
        r->primary = (uint8_t)primary;
        memset(r->pad, 0, sizeof(r->pad));
        memset(r->cigar, 0, sizeof(r->cigar));
        size_t cl = strlen(cigar);
        if (cl >= ORC_CIGAR_TEXT) cl = ORC_CIGAR_TEXT - 1;
        memcpy(r->cigar, cigar, cl);
        r->cigar[cl] = 0;
    }
}

/* put_cigar (mapper_cpe.c:1437-1507): the flush; only the state that changes
 * later results is modelled (block count, per-read record count, reset). */
static void put_cigar(emit_t *em)
{
    *em->read_count += em->cigar_num;
    if (em->cigar_num > 0) {
        if (em->main_flag) em->main_flag = 0;
        em->block++;
    }
    em->cigar_num = 0;
    em->cigar_len = 0;
}

/* Window bit access for align_reads' traceback (REF_LBIT/REF_HBIT,
 * mapper_cpe.c:374-378) at a window-relative index that may leave the window. */
static inline uint32_t win_base(const orc_ref_t *ref, uint32_t win_start, int rfi)
{
    int64_t g = (int64_t)win_start + rfi;
    uint32_t lo = genome_word(ref, g, 0) & 1u, hi = genome_word(ref, g, 1) & 1u;
    return lo | (hi << 1);
}

static inline uint32_t bit_at(uint32_t w, int s) { return (s < 0 || s >= 32) ? 0u : ((w >> s) & 1u); }

/* align_reads for one lane (mapper_cpe.c:243-515). */
static void align_one(const orc_ref_t *ref, emit_t *em, const uint32_t *w, const uint32_t *win_lo,
                      const uint32_t *win_hi, uint32_t ref_len, uint32_t L, uint32_t e,
                      uint32_t win_start, int error, int location)
{
    if (error > (int)e) return;                                             /* :244 */
    int start_location = location - (int)L + 1;
    int tmp_err = 0;
    for (uint32_t ii = 0; ii * 32 < L; ii++) {                              /* :252-276 Hamming */
        int res = 32;
        uint32_t mask = 0xffffffffu;
        if ((ii + 1) * 32 > L) { res = (int)(L - ii * 32); mask >>= (32 - res); }
        uint32_t lref = shr32(win_lo[ii], start_location) | shl32(win_lo[ii + 1], 32 - start_location);
        uint32_t href = shr32(win_hi[ii], start_location) | shl32(win_hi[ii + 1], 32 - start_location);
        uint32_t lread = w[16 + 2 * ii], hread = w[16 + 2 * ii + 1];
        uint32_t x = ((hread ^ href) | (lread ^ lref)) & mask;
        tmp_err += __builtin_popcount(x);
    }
    if (tmp_err == error) {                                                 /* :277-287 */
        uint32_t text_len = 2 + (uint32_t)ndigits(L) + 1;
        char cig[16];
        int cl = sprint_length(cig, L);                                     /* :283-287 "<L>M" */
        cig[cl++] = 'M';
        cig[cl] = 0;
        emit_record(em, (uint32_t)((uint64_t)win_start + (uint64_t)(int64_t)start_location + 1), error, text_len, cig);
        return;
    }
    /* indel path: re-run the recurrence keeping D0/HP per read base (:288-362) */
    uint32_t d0a[ORC_READ_LEN], hpa[ORC_READ_LEN];
    {
        uint32_t band_length = 2 * e + 1, band_down = 2 * e;
        uint32_t peq[4] = {0, 0, 0, 0}, tmp = 1;
        for (uint32_t j = 0; j < band_length; j++) {
            uint32_t b1 = (win_lo[0] >> (j & 0x1f)) & 1u, b2 = (win_hi[0] >> (j & 0x1f)) & 1u;
            peq[b1 | (b2 << 1)] |= tmp;
            tmp <<= 1;
        }
        uint32_t vp = 0, vn = 0, x, d0, hn, hp, i_bd = band_down, mask = 1u << (band_length - 1);
        for (uint32_t ii = 0; ii * 16 < L; ii++) {
            uint32_t res = 16;
            if ((ii + 1) * 16 > L) res = L - ii * 16;
            for (uint32_t i = 0; i < res; i++) {
                uint32_t be = (w[ii] >> (i * 2)) & 3u;
                x = peq[be] | vn;
                d0 = (((x & vp) + vp) ^ vp) | x;
                hn = vp & d0;
                hp = vn | ~(vp | d0);
                x = d0 >> 1;
                vn = x & hp;
                vp = hn | ~(x | hp);
                d0a[ii * 16 + i] = d0;
                hpa[ii * 16 + i] = hp;
                peq[0] >>= 1; peq[1] >>= 1; peq[2] >>= 1; peq[3] >>= 1;
                ++i_bd;
                uint32_t rw = i_bd >> 5;
                uint32_t e1 = (win_lo[rw] >> (i_bd & 0x1f)) & 1u, e2 = (win_hi[rw] >> (i_bd & 0x1f)) & 1u;
                peq[e1 | (e2 << 1)] |= mask;
            }
        }
    }
    int last_high = (int)ref_len - (int)L;
    int match_site = location;
    int site = (int)ref_len - last_high - 1;
    int search_site = match_site - site;
    int pre_size = 1;
    char pre_char = 'N';
    int i = (int)L - 1;
    int sum_err = 0;
    int err = error;
    int rsize[ORC_READ_LEN + 8];
    char rchar[ORC_READ_LEN + 8];
#define RB(rdi) ((w[(rdi) / 16] >> (2 * ((rdi) % 16))) & 3u)
    int j = 1;                                                              /* :379-410 first step */
#if ORC_MUTANT == 5
    if (bit_at(hpa[i], search_site)) {    /* planted mistake: an insertion preferred to a match */
        i--; search_site++; pre_size = 1; pre_char = 'S'; sum_err++; start_location++;
    } else
#endif
    if (bit_at(d0a[i], search_site) && win_base(ref, win_start, match_site) == RB(i)) {
        i--; pre_size = 1; pre_char = 'M'; match_site--;
    } else if (!bit_at(d0a[i], search_site) && win_base(ref, win_start, match_site) != RB(i)) {
        i--; pre_size = 1; pre_char = 'S'; match_site--; sum_err++;
    } else if (bit_at(hpa[i], search_site)) {
#if ORC_MUTANT == 4
        i--; search_site++; pre_size = 1; pre_char = 'I'; sum_err++; start_location++;   /* planted mistake */
#else
        i--; search_site++; pre_size = 1; pre_char = 'S'; sum_err++; start_location++;
#endif
    } else {
        search_site--; pre_size = 1; pre_char = 'D'; match_site--; sum_err++; start_location--;
    }
    while (i >= 0) {                                                        /* :412-466 */
        if (sum_err == err) break;
#if ORC_MUTANT == 5
        if (bit_at(hpa[i], search_site)) {
            i--; search_site++; sum_err++;
            if (pre_char != 'I') { rsize[j] = pre_size; rchar[j++] = pre_char; pre_size = 1; pre_char = 'I'; }
            else pre_size++;
            start_location++;
        } else
#endif
        if (bit_at(d0a[i], search_site) && win_base(ref, win_start, match_site) == RB(i)) {
            i--; match_site--;
            if (pre_char != 'M') { rsize[j] = pre_size; rchar[j++] = pre_char; pre_size = 1; pre_char = 'M'; }
            else pre_size++;
        } else if (!bit_at(d0a[i], search_site) && win_base(ref, win_start, match_site) != RB(i)) {
            i--; match_site--; sum_err++;
            if (pre_char != 'S') { rsize[j] = pre_size; rchar[j++] = pre_char; pre_size = 1; pre_char = 'S'; }
            else pre_size++;
        } else if (bit_at(hpa[i], search_site)) {
            i--; search_site++; sum_err++;
            if (pre_char != 'I') { rsize[j] = pre_size; rchar[j++] = pre_char; pre_size = 1; pre_char = 'I'; }
            else pre_size++;
            start_location++;
        } else {
            search_site--; match_site--; sum_err++;
            if (pre_char != 'D') { rsize[j] = pre_size; rchar[j++] = pre_char; pre_size = 1; pre_char = 'D'; }
            else pre_size++;
            start_location--;
        }
    }
#undef RB
    rsize[j] = pre_size; rchar[j++] = pre_char;                             /* :468-473 */
    if (i >= 0) { rsize[j] = i + 1; rchar[j++] = 'M'; }
    uint64_t new_loc = (uint64_t)win_start + (uint64_t)(int64_t)start_location + 1;
    if (em->cigar_num > 0 && (uint64_t)em->location[em->cigar_num - 1] == new_loc) return; /* :475-477 */
    uint32_t text_len = 2;                                                  /* :479-507 */
    char cig[ORC_CIGAR_TEXT + 16];
    int cl = 0;
    for (j = j - 1; j > 0; j--) {
        if (rchar[j] == 'M' || rchar[j] == 'S') {
            int size_sm = 0;
            while (j > 0 && (rchar[j] == 'M' || rchar[j] == 'S')) { size_sm += rsize[j]; j--; }
            j++;
            text_len += (uint32_t)ndigits((uint32_t)size_sm) + 1;
            if (cl < ORC_CIGAR_TEXT - 12) { cl += sprint_length(cig + cl, (uint32_t)size_sm); cig[cl++] = 'M'; }
        } else {
            text_len += (uint32_t)ndigits((uint32_t)rsize[j]) + 1;
            if (cl < ORC_CIGAR_TEXT - 12) { cl += sprint_length(cig + cl, (uint32_t)rsize[j]); cig[cl++] = rchar[j]; }
        }
    }
    cig[cl] = 0;
    emit_record(em, (uint32_t)new_loc, error, text_len, cig);
}

/* banded_myers (mapper_cpe.c:1590-1864): groups of up to 8 valid lanes, each
 * group verified, then recorded, then the flush test (:1790, :1858). */
static void banded_myers(const orc_ref_t *ref, emit_t *em, const strand_ws_t *ws, uint32_t L,
                         uint32_t e, const uint32_t *cands, uint32_t n_cands, uint32_t batch,
                         uint64_t *n_verif, orc_cand_t **ctrace, uint64_t *n_ct, uint64_t *cap_ct)
{
    const uint32_t nwin = (L + 2 * e + 31) / 32 + 2;
    uint32_t lane_lo[8][16], lane_hi[8][16], lane_start[8], lane_len[8], lane_cand[8], lane_idx[8];
    uint32_t i = 0;
    while (i < n_cands) {
        uint32_t nv = 0;
        while (nv < 8 && i < n_cands) {
            uint64_t shift;
            uint32_t rl;
            uint32_t st = pos_start(ref, e, cands[i], &shift, &rl, L);
            uint32_t idx = i;
            i++;
            if (st == 0xffffffffu) continue;
            for (uint32_t q = 0; q < nwin; q++) {
                lane_lo[nv][q] = genome_word(ref, (int64_t)st + 32 * (int64_t)q, 0);
                lane_hi[nv][q] = genome_word(ref, (int64_t)st + 32 * (int64_t)q, 1);
            }
            lane_start[nv] = st; lane_len[nv] = rl; lane_cand[nv] = cands[idx]; lane_idx[nv] = idx;
            nv++;
        }
        int32_t er[8], lc[8];
        for (uint32_t q = 0; q < nv; q++) {
            orc_myers(ws->w, lane_lo[q], lane_hi[q], L, lane_len[q], e, &er[q], &lc[q]);
            if (n_verif) (*n_verif)++;
            if (ctrace) {
                if (*n_ct == *cap_ct) {
                    *cap_ct = *cap_ct ? *cap_ct * 2 : 256;
                    *ctrace = (orc_cand_t *)realloc(*ctrace, *cap_ct * sizeof(orc_cand_t));
                }
                orc_cand_t *c = &(*ctrace)[(*n_ct)++];
                c->read = em->read; c->strand = (uint8_t)em->strand; c->pad0 = 0;
                c->batch = (uint16_t)batch; c->idx = lane_idx[q]; c->cand = lane_cand[q];
                c->win_start = lane_start[q]; c->ref_len = lane_len[q]; c->err = er[q]; c->loc = lc[q];
            }
        }
        for (uint32_t q = 0; q < nv; q++)
            align_one(ref, em, ws->w, lane_lo[q], lane_hi[q], lane_len[q], L, e, lane_start[q], er[q], lc[q]);
        if (em->cigar_num + 8 > ORC_MAX_CIGAR || em->cigar_len + 16 * ORC_CIGAR_LEN > ORC_MAX_CIGAR * ORC_CIGAR_LEN)
            put_cigar(em);
    }
}

/* ---- candidate batches (mapper_cpe.c:1893-2123) ---- */

typedef struct { uint32_t loc; uint32_t seed_ind; } sh_t;  /* seed_ind_t :109-112 */
typedef struct { uint32_t pos; uint32_t pos_ind; } ph_t;   /* pos_heap_t :1995-1998 */

/* adjust_ref_pos_rmdup (mapper_cpe.c:1893-1917); the "ajr_flag = 1"
 * assignment at :1908 makes every nonzero flag take the second branch. */
static void adjust_ref_pos(uint32_t k, uint32_t *positions, const hc_t *info, const uint32_t *pos_offset,
                           uint64_t pre_i, uint32_t ajr_flag, uint64_t num)
{
    if (ajr_flag == 0) {
        uint64_t pos_off = 0;
        for (uint64_t i = 0; i < num; i++) {
            uint64_t base_offset = (uint64_t)info[i / ORC_SEED_STEP].index * (k + ORC_SEED_STEP - 1) + i % ORC_SEED_STEP;
            for (uint64_t j = pos_off; j < pos_offset[i]; j++)
                positions[j] = positions[j] > base_offset ? (uint32_t)(positions[j] - base_offset) : 0;
            pos_off = pos_offset[i];
        }
    } else {
        uint64_t i = pre_i;
        uint64_t base_offset = (uint64_t)info[i / ORC_SEED_STEP].index * (k + ORC_SEED_STEP - 1) + i % ORC_SEED_STEP;
        for (uint64_t j = 0; j < pos_offset[0]; j++)
            positions[j] = positions[j] > base_offset ? (uint32_t)(positions[j] - base_offset) : 0;
    }
}

static void seeds_heap_insert(sh_t *h, uint32_t *size, uint32_t pos, uint32_t seed_ind)   /* :1919-1935 */
{
    *size += 1;
    h[*size].loc = pos;
    h[*size].seed_ind = seed_ind;
    uint32_t i = *size;
    while (i > 1) {
        uint32_t ti = i / 2;
        if (h[ti].loc > h[i].loc) { sh_t t = h[ti]; h[ti] = h[i]; h[i] = t; }
        else break;
        i = ti;
    }
}

static void seeds_heap_pop_and_insert(sh_t *h, uint32_t *size, uint32_t pos, uint32_t del_flag) /* :1936-1962 (Q6) */
{
    if (del_flag == 1) { h[1] = h[*size]; *size -= 1; }
    else h[1].loc = pos;
    uint32_t i = 1, ti;
#if ORC_MUTANT == 1
    while (i * 2 <= *size) {            /* planted mistake: the textbook sift-down */
#else
    while (i * 2 < *size) {
#endif
        ti = i * 2;
        if (i * 2 + 1 <= *size)
            if (h[i * 2].loc > h[i * 2 + 1].loc) ti = i * 2 + 1;
        if (h[ti].loc < h[i].loc) { sh_t t = h[ti]; h[ti] = h[i]; h[i] = t; }
        else break;
        i = ti;
    }
}

/* block_heap_sort (mapper_cpe.c:1999-2089): k-way merge of the ascending
 * blocks with neighbour dedup into sorted[]; count returned in pos_ind[0]. */
static void block_heap_sort(const uint32_t *positions, uint32_t *sorted, ph_t *heap, uint32_t *pos_ind,
                            uint32_t block_size)
{
    uint32_t heap_size = 0;
    for (uint32_t i = 0; i < block_size / 2; i++) {
        if (pos_ind[i * 2] < pos_ind[2 * i + 1]) {
            heap[++heap_size].pos = positions[pos_ind[2 * i]];
            heap[heap_size].pos_ind = 2 * i;
            pos_ind[2 * i]++;
            uint64_t cur = heap_size, par = cur / 2;
            while (cur > 1) {
                if (heap[par].pos > heap[cur].pos) { ph_t t = heap[cur]; heap[cur] = heap[par]; heap[par] = t; }
                else break;
                cur /= 2;
                par = cur / 2;
            }
        }
    }
    if (heap_size == 0) { pos_ind[0] = 0; return; }                          /* Q8 */
    uint32_t pre_small = 0xffffffffu, cur_ind = 0;
    while (heap_size > 1) {
        uint32_t pos_id = heap[1].pos_ind;
        if (pre_small != heap[1].pos) { sorted[cur_ind++] = heap[1].pos; pre_small = heap[1].pos; }
        if (pos_ind[pos_id] < pos_ind[pos_id + 1]) {
            heap[1].pos = positions[pos_ind[pos_id]];
            pos_ind[pos_id]++;
        } else {
            heap[1] = heap[heap_size];
            heap_size--;
            if (heap_size == 1) break;
        }
        uint32_t hi = 1;
        while (hi * 2 <= heap_size) {
            uint32_t l = hi * 2, r = l + 1, s = l;
            if (r <= heap_size && heap[l].pos > heap[r].pos) s = r;
            if (heap[hi].pos > heap[s].pos) { ph_t t = heap[hi]; heap[hi] = heap[s]; heap[s] = t; }
            else break;
            hi = s;
        }
    }
    uint32_t pos_id = heap[1].pos_ind;
    for (uint32_t i = pos_ind[pos_id] - 1; i < pos_ind[pos_id + 1]; i++) {
        uint32_t p = positions[i];
        if (p != pre_small) { sorted[cur_ind++] = p; pre_small = p; }
    }
    pos_ind[0] = cur_ind;
}

/* long_seeds_filter (mapper_cpe.c:2090-2123), including the drop at :2113 (Q7). */
static void long_seeds_filter(const strand_ws_t *ws, uint32_t *positions, const uint16_t *lseeds,
                              uint32_t *pos_ind, const uint16_t *lseed_ind, const uint16_t *lseed_off,
                              uint32_t block_size)
{
    const uint32_t res_len = ORC_SEED_STEP - 1, mask = (1u << res_len) - 1;
    uint32_t pos_size = 0;
    for (uint32_t i = 0; i < block_size / 2; i++) {
        uint32_t lind = lseed_ind[i];
        uint32_t shift = ORC_SEED_STEP - 1 - lseed_ind[i] % ORC_SEED_STEP;
        uint32_t off = lseed_off[i];
        uint32_t rl = ws->res[2 * lind], rh = ws->res[2 * lind + 1];
        for (uint32_t j = pos_ind[i * 2]; j < pos_ind[2 * i + 1]; j++) {
            uint32_t sl = (uint32_t)lseeds[off] >> shift;
            uint32_t shh = (uint32_t)lseeds[off] >> (res_len * 2 + shift);
            uint32_t r = ((sl ^ rl) | (shh ^ rh)) & mask;
            if (r == 0 && pos_size != j) positions[pos_size++] = positions[j];
            off++;
        }
        if (i > 0) pos_ind[i * 2] = pos_ind[i * 2 - 1];
        pos_ind[i * 2 + 1] = pos_size;
    }
}

/* The seed heap's key of a fetch by 12-mer i: its raw last position (Q6:
 * the reference keys the heap before adjust_ref_pos_rmdup shifts it). */
static inline uint32_t heap_key(uint32_t k, const hc_t *info, uint64_t i, uint32_t raw)
{
#if ORC_MUTANT == 2
    /* planted mistake: the key shifted to the read start */
    uint64_t off = (uint64_t)info[i / ORC_SEED_STEP].index * (k + ORC_SEED_STEP - 1) + i % ORC_SEED_STEP;
    return raw > off ? (uint32_t)(raw - off) : 0u;
#else
    (void)k; (void)info; (void)i;
    return raw;
#endif
}

/* alignment_rmdup (mapper_cpe.c:2124-2426) for one strand. */
static void alignment_rmdup(const orc_ref_t *ref, const strand_ws_t *ws, uint32_t L, uint32_t e,
                            emit_t *em, uint64_t *n_verif,
                            orc_cand_t **ctrace, uint64_t *n_ct, uint64_t *cap_ct)
{
    const uint32_t k = ref->k;
    const uint64_t positions_size = 2048, per_block = 32;
    uint32_t positions[2048 + 64];
    uint32_t sorted[2048 + 64];
    uint16_t lseeds[4096 + 64];
    uint32_t pos_ind[264];
    uint16_t lseed_ind[128 + 8], lseed_off[128 + 8];
    ph_t pheap[264];
    sh_t sheap[64 + 2];
    uint32_t sheap_size = 0;
    uint64_t pos_count[64];
    uint32_t pos_offset[64];
    uint32_t block_size = 1;
    uint32_t pos_size = 0, lseed_size = 0, lseed_get = 0, lseed_o = 0;
    uint32_t batch = 0;

    uint64_t n_long = ws->n_seeds / ORC_SEED_STEP;
    uint64_t n_sel = (uint64_t)e + 1 < n_long ? (uint64_t)e + 1 : n_long;  /* Q5 */
    uint64_t seeds_num = n_sel * ORC_SEED_STEP;

    em->main_flag = 1;
    em->cigar_len = 0;
    em->cigar_num = 0;
    em->block = 0;

    pos_ind[0] = 0;
    pos_ind[1] = 0;
    for (uint64_t i = 0; i < seeds_num; i++) {                              /* :2194-2236 phase 1 */
        uint64_t si = (uint64_t)ws->info[i / ORC_SEED_STEP].index * ORC_SEED_STEP + i % ORC_SEED_STEP;
        uint64_t cnt = ws->cnt[si];
        uint64_t pos_get = cnt > per_block ? per_block : cnt;
        if (pos_get > 0) {
            uint32_t off = ws->off2[si * 2];
            memcpy(positions + pos_size, ref->positions + off, pos_get * sizeof(uint32_t));
            lseed_o = off - (off & 1u);
            lseed_get = (uint32_t)pos_get + (off & 1u);
            lseed_get = lseed_get + (lseed_get & 1u);
            for (uint32_t q = 0; q < lseed_get; q++)
                lseeds[lseed_size + q] = (uint64_t)lseed_o + q < ref->n_pos ? ref->pos_lseeds[lseed_o + q] : 0;
        }
        pos_size += (uint32_t)pos_get;
        pos_count[i] = pos_get;
        pos_offset[i] = pos_size;
        if (pos_get > 0) {
            uint32_t off = ws->off2[si * 2];
            lseed_off[block_size / 2] = (uint16_t)(lseed_size + (off & 1u));
            lseed_ind[block_size / 2] = (uint16_t)si;
            lseed_size += lseed_get;
            pos_ind[block_size++] = pos_size;
            pos_ind[block_size++] = pos_size;
        }
    }
    for (uint64_t i = 0; i < seeds_num; i++) {                              /* :2244-2252 */
        uint64_t si = (uint64_t)ws->info[i / ORC_SEED_STEP].index * ORC_SEED_STEP + i % ORC_SEED_STEP;
        if (pos_count[i] != ws->cnt[si])
            seeds_heap_insert(sheap, &sheap_size, heap_key(k, ws->info, i, positions[pos_offset[i] - 1]), (uint32_t)i);
    }
    uint32_t pre_pos_size = 0, ajr_flag = 0;
    uint64_t pre_i = 0;
    for (;;) {                                                              /* :2261-2419 */
#if ORC_MUTANT == 3
        if (sheap_size > 0 && pos_size + per_block < positions_size && block_size < 256) {   /* planted mistake */
#else
        if (sheap_size > 0 && pos_size + per_block <= positions_size && block_size < 256) {
#endif
            uint64_t i = sheap[1].seed_ind;
            uint64_t si = (uint64_t)ws->info[i / ORC_SEED_STEP].index * ORC_SEED_STEP + i % ORC_SEED_STEP;
            uint64_t cnt = ws->cnt[si];
            uint64_t pos_get = cnt > pos_count[i] + per_block ? per_block : cnt - pos_count[i];
            uint32_t lo_odd = 0;
            if (pos_get > 0) {
                uint32_t off = ws->off2[si * 2] + (uint32_t)pos_count[i];
                memcpy(positions + pos_size, ref->positions + off, pos_get * sizeof(uint32_t));
                lseed_o = off - (off & 1u);
                lseed_get = (uint32_t)pos_get + (off & 1u);
                lseed_get = lseed_get + (lseed_get & 1u);
                for (uint32_t q = 0; q < lseed_get; q++)
                    lseeds[lseed_size + q] = (uint64_t)lseed_o + q < ref->n_pos ? ref->pos_lseeds[lseed_o + q] : 0;
                lo_odd = off & 1u;
            }
            if (pos_size > 0)
                adjust_ref_pos(k, positions + pre_pos_size, ws->info, pos_offset, pre_i, ajr_flag, seeds_num);
            pre_pos_size = pos_size;
            pos_offset[0] = (uint32_t)pos_get;
            pre_i = i;
            ajr_flag = 1;
            pos_size += (uint32_t)pos_get;
            pos_count[i] += pos_get;
            if (pos_get > 0) {
                lseed_off[block_size / 2] = (uint16_t)(lseed_size + lo_odd);
                lseed_ind[block_size / 2] = (uint16_t)si;
                lseed_size += lseed_get;
                pos_ind[block_size++] = pos_size;
                pos_ind[block_size++] = pos_size;
            }
            uint32_t del_flag = pos_count[i] == cnt ? 1u : 0u;
            seeds_heap_pop_and_insert(sheap, &sheap_size, heap_key(k, ws->info, i, positions[pos_size - 1]), del_flag);
        }
#if ORC_MUTANT == 3
        if (pos_size + per_block >= positions_size || sheap_size == 0 || block_size >= 256) {
#else
        if (pos_size + per_block > positions_size || sheap_size == 0 || block_size >= 256) {
#endif
            adjust_ref_pos(k, positions + pre_pos_size, ws->info, pos_offset, pre_i, ajr_flag, seeds_num);
            if (pos_size > 0) {
                long_seeds_filter(ws, positions, lseeds, pos_ind, lseed_ind, lseed_off, block_size - 1);
                lseed_size = 0;
                pos_size = pos_ind[block_size - 2];
                block_heap_sort(positions, sorted, pheap, pos_ind, block_size - 1);
                block_size = 1;
                pos_size = pos_ind[0];
            }
            if (pos_size > 0)
                banded_myers(ref, em, ws, L, e, sorted, pos_size, batch, n_verif, ctrace, n_ct, cap_ct);
            batch++;
            pos_size = 0;
            pos_offset[0] = 0;
            pre_pos_size = 0;
            ajr_flag = 2;
            block_size = 1;
            pos_ind[0] = 0;
            pos_ind[1] = 0;
            if (sheap_size == 0) break;
        }
    }
    put_cigar(em);                                                          /* :2421-2424 */
}

/* get_positions for one read (mapper_cpe.c:2500-2590). */
static void map_one(const orc_ref_t *ref, const char *read, uint32_t L, uint32_t e, uint32_t read_id,
                    orc_hit_t *hit, uint64_t *n_verif, int keep,
                    orc_record_t **recs, uint64_t *n_recs, uint64_t *cap_recs,
                    orc_cand_t **ctrace, uint64_t *n_ct, uint64_t *cap_ct)
{
    strand_ws_t ws;
    char rev[ORC_READ_LEN + 8];
    uint32_t read_count = 0;
    emit_t em;
    memset(&em, 0, sizeof(em));
    em.read = read_id;
    em.read_count = &read_count;
    em.hit = hit;
    em.keep_records = keep;
    if (keep) { em.recs = *recs; em.n_recs = *n_recs; em.cap_recs = *cap_recs; }
    hit->pos = 0xffffffffu; hit->strand = 0xff; hit->dist = 0xff; hit->pad = 0; hit->count = 0;

    /* more than 16 long seeds overflow the reference's per-read tables
     * (hash_seed_info[16], hash_count[64], mapper_cpe.c:2503-2515): such
     * reads are defined unmapped */
    if (L / (ref->k + ORC_SEED_STEP - 1) > 16) L = 0;
    for (int strand = 0; strand < 2; strand++) {
        const char *rs = read;
        if (strand == 1) { orc_reverse_read(rev, read, L); rs = rev; }
        memset(&ws, 0, sizeof(ws));
        rd_choose_seed(ref, rs, L, &ws);
        rd_long_seeds(ref->k, L, &ws);
        em.strand = (uint32_t)strand;
        alignment_rmdup(ref, &ws, L, e, &em, n_verif, ctrace, n_ct, cap_ct);
    }
    if (hit->count == 0) { hit->pos = 0xffffffffu; hit->strand = 0xff; hit->dist = 0xff; }
    if (keep) { *recs = em.recs; *n_recs = em.n_recs; *cap_recs = em.cap_recs; }
}

static int check_args(const orc_ref_t *ref, uint64_t stride, uint32_t e)
{
    if (!ref || !ref->hash_offset || !ref->base_2bits) return -1;
    if (e > 15 || stride == 0) return -1;
    return 0;
}

int orc_map(const orc_ref_t *ref, const char *reads, uint64_t stride, const uint32_t *lens,
            uint64_t n_reads, uint32_t e, orc_hit_t *hits, uint64_t *n_verif, int threads)
{
    if (check_args(ref, stride, e)) return -1;
    uint64_t total = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
#endif
    for (int64_t r = 0; r < (int64_t)n_reads; r++) {
        uint64_t nv = 0;
        uint32_t L = lens[r] > ORC_READ_LEN - 2 ? ORC_READ_LEN - 2 : lens[r];
        map_one(ref, reads + (uint64_t)r * stride, L, e, (uint32_t)r, &hits[r], &nv, 0,
                NULL, NULL, NULL, NULL, NULL, NULL);
        total += nv;
    }
    if (n_verif) *n_verif = total;
    return 0;
}

int orc_map_trace(const orc_ref_t *ref, const char *reads, uint64_t stride, const uint32_t *lens,
                  uint64_t n_reads, uint32_t e, orc_hit_t *hits,
                  orc_record_t **recs, uint64_t *n_recs, orc_cand_t **cands, uint64_t *n_cands)
{
    if (check_args(ref, stride, e)) return -1;
    orc_record_t *rv = NULL; uint64_t nr = 0, cr = 0;
    orc_cand_t *cv = NULL; uint64_t nc = 0, cc = 0;
    for (uint64_t r = 0; r < n_reads; r++) {
        uint32_t L = lens[r] > ORC_READ_LEN - 2 ? ORC_READ_LEN - 2 : lens[r];
        map_one(ref, reads + r * stride, L, e, (uint32_t)r, &hits[r], NULL, 1, &rv, &nr, &cr,
                cands ? &cv : NULL, &nc, &cc);
    }
    if (recs) *recs = rv; else free(rv);
    if (n_recs) *n_recs = nr;
    if (cands) *cands = cv; else free(cv);
    if (n_cands) *n_cands = nc;
    return 0;
}

void orc_free(void *p) { free(p); }
