/*
 * swm_oracle.h -- CPU restatement of SWMapper's mapping hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is compared
 * against; it is never linked into the product library.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every function restates one function of the reference (Xu-Kai/SWMapper) and
 * cites the file:line it follows.  Where the reference is undefined (shift by
 * 32, stale local buffers, uninitialised malloc memory) the oracle fixes the
 * semantics named in the comment at that spot ("define" entries Q1..Q23).
 */
#ifndef SWM_ORACLE_H
#define SWM_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_SEED_STEP 4          /* fm_define.h:13 */
#define ORC_MAXERROR 10000       /* mapper_cpe.c:17 */
#define ORC_MAX_CIGAR 256        /* fm_define.h:17 */
#define ORC_CIGAR_LEN 16         /* fm_define.h:19 */
#define ORC_READ_LEN 256         /* mapper_cpe.c:15 */

/* ---- index (serial_index.c, HOST_BUILD) ---- */
typedef struct orc_index {
    uint32_t k;
    uint64_t n_bases;          /* bases over all contigs ("position" at the end) */
    uint64_t n_words;          /* ceil(n_bases / 32) */
    uint32_t *base_2bits;      /* 2 * n_words words, file .base_2bits */
    uint32_t *base_3bits;      /* n_words words, file .base_3bits */
    uint32_t *hash_count;      /* 4^k words, file .hash_count */
    uint32_t *hash_offset_end; /* 4^k + 1 words, file .hash_offset (END offsets) */
    uint32_t *positions;       /* n_pos words, file .positions */
    uint16_t *pos_lseeds;      /* n_pos half-words, file .pos_lseeds */
    uint64_t n_pos;
} orc_index_t;

/* seq: the contigs concatenated in FASTA order (ASCII, any case). */
int  orc_index_build(const char *seq, uint64_t n, uint32_t k, orc_index_t *out);
void orc_index_free(orc_index_t *ix);

/* ---- mapper (mapper_cpe.c) ---- */
typedef struct orc_ref {
    const uint32_t *hash_offset;  /* exclusive scan of .hash_count, 4^k + 1 (mapper_mpe.c:250-255) */
    const uint32_t *positions;
    const uint16_t *pos_lseeds;
    const uint32_t *base_2bits;   /* file words followed by >= 64 zero words (Q2) */
    uint64_t n_2bits_words;       /* words in the .base_2bits file */
    uint64_t ref_length;          /* 16 * bytes(.base_2bits) (mapper_mpe.c:288) */
    uint64_t n_pos;
    uint32_t k;
} orc_ref_t;

/* Per-read result (the four outputs the build reports per read). */
typedef struct orc_hit {
    uint32_t pos;     /* 1-based genome coordinate of the best record, UINT32_MAX if unmapped */
    uint8_t  strand;  /* 0 forward, 1 reverse complement, 0xFF if unmapped */
    uint8_t  dist;    /* edit distance of the best record, 0xFF if unmapped */
    uint16_t pad;
    uint32_t count;   /* records emitted over both strands (duplicates included) */
} orc_hit_t;

/* One emitted record, in emission order. */
#define ORC_CIGAR_TEXT 160
typedef struct orc_record {
    uint32_t read;
    uint8_t  strand;
    uint8_t  dist;
    uint16_t block;   /* flush block number within (read, strand) */
    uint32_t loc;     /* cigar->location value: window_start + start_location + 1 */
    uint32_t text_len;/* bytes this record adds to the CIGAR buffer */
    uint8_t  primary; /* the first record of its (read, strand): main_flag (mapper_cpe.c:1462) */
    uint8_t  pad[3];
    char     cigar[ORC_CIGAR_TEXT];  /* the CIGAR text (align_reads :283-287, :479-507), NUL-terminated */
} orc_record_t;

/* One verification (one valid Myers lane), in evaluation order. */
typedef struct orc_cand {
    uint32_t read;
    uint8_t  strand;
    uint8_t  pad0;
    uint16_t batch;     /* candidate batch within (read, strand) */
    uint32_t idx;       /* index within the batch's sorted unique list */
    uint32_t cand;      /* adjusted read-start candidate (positions[2048+idx]) */
    uint32_t win_start; /* pos_start(): cand - shift */
    uint32_t ref_len;
    int32_t  err;
    int32_t  loc;
} orc_cand_t;

/* Map reads[i*stride .. +lens[i]) for i < n_reads.  hits[n_reads] required.
 * n_verif (optional) receives the number of Myers verifications.
 * threads <= 0 uses the OpenMP default. */
int orc_map(const orc_ref_t *ref, const char *reads, uint64_t stride,
            const uint32_t *lens, uint64_t n_reads, uint32_t e,
            orc_hit_t *hits, uint64_t *n_verif, int threads);

/* Single-threaded map that also returns the ordered records and candidates.
 * Arrays are malloc'ed; free with orc_free(). */
int orc_map_trace(const orc_ref_t *ref, const char *reads, uint64_t stride,
                  const uint32_t *lens, uint64_t n_reads, uint32_t e,
                  orc_hit_t *hits,
                  orc_record_t **recs, uint64_t *n_recs,
                  orc_cand_t **cands, uint64_t *n_cands);
void orc_free(void *p);

/* ---- stage-level entry points for per-kernel parity tests ---- */

/* Read encoding of one strand (get_seed_from_reads over all seeds).
 * words32[32]: [0..15] interleaved 2-bit words, [16..31] bit planes.
 * seeds[64]: seed code of seed j.  Returns the number of seeds. */
int orc_encode_read(const char *read, uint32_t L, uint32_t k,
                    uint32_t *words32, uint32_t *seeds);

/* Reverse complement as the reference forms it (mapper_cpe.c:2429-2450). */
void orc_reverse_read(char *out, const char *read, uint32_t L);

/* Banded Myers on one candidate.  read_words: the 16 interleaved words;
 * win_lo/win_hi: window planes aligned at the window start, >= ceil((L+2e)/32)+1
 * words each.  Returns err (MAXERROR if > e) and loc via pointers. */
void orc_myers(const uint32_t *read_words, const uint32_t *win_lo,
               const uint32_t *win_hi, uint32_t L, uint32_t ref_len, uint32_t e,
               int32_t *err, int32_t *loc);

/* Same recurrence with a 64-bit band word (Q14 width check). */
void orc_myers64(const uint32_t *read_words, const uint32_t *win_lo,
                 const uint32_t *win_hi, uint32_t L, uint32_t ref_len, uint32_t e,
                 int32_t *err, int32_t *loc);

#ifdef __cplusplus
}
#endif
#endif
