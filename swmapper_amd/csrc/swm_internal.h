/*
 * swm_internal.h -- shared definitions of the HIP mapping path.
 */
#ifndef SWM_INTERNAL_H
#define SWM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

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

/* Per (read, strand) seed selection entry, 8 bytes: the selected 12-mer's
 * bucket start in the device bucket array, and cm = count << 16 | meta.
 * meta:
 *   bits  0..7   adjust offset 15*m + q (< the read length; adjust_ref_pos_rmdup, mapper_cpe.c:1901)
 *   bits  8..10  residual, low plane  (set_read_long_seeds, mapper_cpe.c:2489)
 *   bits 11..13  residual, high plane (mapper_cpe.c:2494)
 *   bits 14..15  q = seed_index % 4 (long_seeds_filter shift, mapper_cpe.c:2101)
 * A count of 0xffff or more is stored as 0xffff and in full in the run's
 * big-count array at the same index (written only for those). */
struct swm_sel_t {
    uint32_t off;
    uint32_t cm;
};
#define SEL_BIG 0xffffu

struct swm_ws;    /* device workspace of one slot */
struct swm_slot;  /* stream + workspace + captured graphs */

/* Order of the slots' pipeline runs on the device: each run waits for the
 * previous one (whatever its slot) before its kernels start, so runs do not
 * share the GPU -- only one slot's copies overlap another's kernels. */
struct swm_gate {
    std::mutex mu;
    hipEvent_t last = nullptr;   /* the previous run's end (an event of its slot) */
};

/* Forms and tunables of the pipeline (swm_ctx_set_option). */
struct swm_opts {
    uint32_t work_queue = 1;        /* 0: the plain one-thread-per-item kernels for every item */
    uint32_t graph = 0;             /* 1: fixed-capacity pipeline replayed as a hipGraph */
    uint32_t heavy_cands = 64;      /* items with more candidates go to the cooperative kernels */
    uint32_t myers_block = 256;     /* workgroup size of the verification kernel */
    uint32_t plan_heavy_waves = 32; /* persistent waves per CU draining the heavy plan queue */
    uint32_t emit_heavy_waves = 32; /* waves per CU of the cooperative emission */
    uint32_t histogram = 0;         /* collect the per-read candidate histogram */
    uint32_t myers_packed = 1;      /* e <= 7: two candidates per lane in 16-bit halves */
    uint32_t giant_pos = 16384;     /* heavy plan items with more positions are taken first */
    uint32_t giant_cands = 16384;   /* heavy emission reads with more candidates are taken first */
    uint32_t seed_generic = 0;      /* 1: the seed kernel for any k even where a fixed-k one exists */
    uint32_t plan_overlap = 0;      /* 1: heavy-item plan kernels on the slot's second stream (work queue) */
    uint32_t plan_light_waves = 32; /* waves per CU of the light-item plan kernel */
};

struct swm_ctx {
    int device = 0;
    int n_cu = 256;                     /* compute units: sizes the persistent work-queue grids */
    uint32_t k = 12;
    uint64_t n_buckets = 0;
    uint2 *d_seed_tab = nullptr;        /* (offset, count) of every k-mer in two roles: two
                                         * consecutive k-mers and their reverse complements in
                                         * one line (8 or 16 x 4^(k-1) entries) */
    uint2 *d_pl = nullptr;              /* bucket entries: (position, 12-bit residual) */
    uint64_t n_pos = 0;
    uint64_t n_dev_pos = 0;     /* entries of the device bucket array (pairs packed into lines) */
    uint32_t *d_base_2bits = nullptr;   /* file words + SWM_WIN_PAIRS_PAD zero pairs */
    uint64_t n_2bits_words = 0;
    uint64_t ref_length = 0;            /* 16 * bytes(.base_2bits) (mapper_mpe.c:288) */
    swm_recipe_t *d_recipe = nullptr;   /* read-encoding recipe per read length 0..254 */
    uint32_t *d_irreg = nullptr;        /* start offsets of the buckets not in ascending order */
    uint32_t n_irreg = 0;
    bool irr_flags = false;             /* bit 31 of d_hash_offset[bucket] flags the irregular ones */
    swm_opts opt;
    swm_slot *slot0 = nullptr;          /* the slot of the synchronous entry points */
    uint8_t *d_scratch = nullptr;       /* this device's no-allocation scratch (SWM_NOALLOC_SCRATCH_BYTES) */
};

swm_slot *swm_slot_create(swm_ctx_t *ctx);
void swm_slot_destroy(swm_slot *s);
const swm_stats_t *swm_slot_stats(const swm_slot *s);
uint64_t swm_slot_bytes(const swm_slot *s);
/* size the slot's workspace for batches of up to n reads once graphs are used */
void swm_slot_reserve(swm_slot *s, uint64_t n);
/* runs of this slot take their turn through the gate (NULL: no gate) */
void swm_slot_set_gate(swm_slot *s, swm_gate *g);
/* map n reads already on the device, on the slot's stream, in the synced or
 * the graph form (returns after the results are written) */
int swm_slot_run_device(swm_ctx_t *ctx, swm_slot *sl, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                        uint64_t n, uint32_t e, swm_read_result_t *d_out, bool graph);
/* map n host reads through the slot's buffers, padded to n_pad reads of length 0 */
int swm_slot_run_host(swm_ctx_t *ctx, swm_slot *sl, const char *reads, uint64_t stride, const uint32_t *lens,
                      uint64_t n, uint64_t n_pad, uint32_t e, swm_read_result_t *out, bool graph);
/* several host batches mapped in one run (results per segment) */
struct swm_segment_t {
    const char *reads;
    const uint32_t *lens;
    uint64_t n;
    swm_read_result_t *out;
};
int swm_slot_run_segments(swm_ctx_t *ctx, swm_slot *sl, const swm_segment_t *seg, uint32_t nseg, uint64_t stride,
                          uint64_t n_pad, uint32_t e, bool graph);
/* SWM_EINVAL if a read breaks the stride rule or has more than 16 long seeds */
int swm_check_lens(const swm_ctx_t *ctx, uint64_t stride, const uint32_t *lens, uint64_t n);

#define SWM_HIP_CHECK(call)                                   \
    do {                                                      \
        hipError_t _e = (call);                               \
        if (_e != hipSuccess) return SWM_EDEVICE;             \
    } while (0)

#endif
