/*
 * swm_internal.h -- shared definitions of the HIP mapping path.
 */
#ifndef SWM_INTERNAL_H
#define SWM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swm.h"
#include "swm_recipe.h"

#define SWM_SEED_STEP 4        /* fm_define.h:13 */
#define SWM_MAXERROR 10000     /* mapper_cpe.c:17 */
#define SWM_READ_MAX 254       /* READ_LEN 256 less the kseq terminator pair (mapper_cpe.c:15) */
#define SWM_MAX_E 15           /* band 2e+1 <= 31 bits (mapper_cpe.c:686) */
#define SWM_BATCH_POS 2048     /* positions per candidate batch (mapper_cpe.c:2129) */
#define SWM_CHUNK 32           /* positions fetched per seed per step (mapper_cpe.c:2178) */
#define SWM_MAX_BLOCKS 128     /* block_size < 256 (mapper_cpe.c:2265,2356) */
#define SWM_WIN_PAIRS_PAD 64   /* zero word pairs after the genome (windows past the end read 0) */

/* Per (read, strand) seed selection entry: packed metadata.
 *   bits  0..15  adjust offset 15*m + q (adjust_ref_pos_rmdup, mapper_cpe.c:1901)
 *   bits 16..18  residual, low plane  (set_read_long_seeds, mapper_cpe.c:2489)
 *   bits 19..21  residual, high plane (mapper_cpe.c:2494)
 *   bits 22..23  q = seed_index % 4 (long_seeds_filter shift, mapper_cpe.c:2101) */
struct swm_sel_t {
    uint32_t off;
    uint32_t cnt;
    uint32_t meta;
};

struct swm_ws;  /* per-context device workspace */

struct swm_ctx {
    int device;
    int n_cu;                   /* compute units: sizes the persistent work-queue grids */
    hipStream_t stream;
    uint32_t k;
    uint64_t n_buckets;
    uint32_t *d_hash_offset;    /* exclusive scan of hash_count, 4^k + 1 */
    uint2 *d_pl;                /* bucket entries: (position, 12-bit residual) */
    uint64_t n_pos;
    uint32_t *d_base_2bits;     /* file words + SWM_WIN_PAIRS_PAD zero pairs */
    uint64_t n_2bits_words;
    uint64_t ref_length;        /* 16 * bytes(.base_2bits) (mapper_mpe.c:288) */
    swm_recipe_t *d_recipe;     /* read-encoding recipe per read length 0..254 */
    swm_stats_t stats;
    swm_ws *ws;
};

#define SWM_HIP_CHECK(call)                                   \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return SWM_EDEVICE;             \
    } while (0)

#endif
