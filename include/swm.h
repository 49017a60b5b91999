/*
 * swm.h -- C ABI of the MI355X mapping hot path (SWMapper drop-in).
 *
 * The reference's boundary for this path is
 *   - the index files written by serial_index.c (serial_index.c:73-138,393-408)
 *     and read by load_ref_info (mapper_mpe.c:198-314);
 *   - athread_spawn(find_positions, &ref_info_t) / athread_join()
 *     (mapper_mpe.c:972,1036,1056,1088; kernel mapper_cpe.c:2592-2663),
 *     with the parameter block ref_info_t (fm_define.h:22-55);
 *   - the mapper options -l -e -s -p (mapper_mpe.c:717-760) and the index
 *     builder options -i -p -s (serial_index.c:819-846).
 * Each entry point below names the reference interface it replaces.
 *
 * All functions return SWM_OK (0) or a negative SWM_E* status.  No function
 * prints or aborts.  A context is safe to use from one host thread at a time;
 * batches submitted on one context run on that context's stream in order.
 */
#ifndef SWM_H
#define SWM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWM_OK 0
#define SWM_EINVAL -1      /* bad argument (option out of range, NULL pointer) */
#define SWM_EIO -2         /* index or sequence file missing, short or mis-sized */
#define SWM_ENOMEM -3      /* host or device allocation failed */
#define SWM_EDEVICE -4     /* HIP runtime error or no device */
#define SWM_EBUSY -5       /* serving loop: no admission possible right now */
#define SWM_ENOSPC -6      /* no-allocation entry: the batch does not fit the fixed scratch */

/* Per-read result: the four per-read outputs of the mapping path.
 * count = records emitted over both strands (the reference's cigar records,
 * duplicates included); best = minimum edit distance, first in emission order
 * (forward strand first); pos = 1-based genome coordinate of that record
 * (cigar->location, mapper_cpe.c:286,484).  Unmapped: count 0, pos
 * UINT32_MAX, strand 0xFF, dist 0xFF. */
typedef struct swm_read_result {
    uint32_t pos;
    uint8_t strand;
    uint8_t dist;
    uint16_t pad;
    uint32_t count;
} swm_read_result_t;

/* In-memory index, the exact content of the reference's index files
 * (.base_2bits, .base_3bits, .hash_count, .hash_offset, .positions,
 * .pos_lseeds).  Arrays are owned by the library when produced by it. */
typedef struct swm_index_host {
    uint32_t k;
    uint64_t n_bases;
    uint64_t n_words;          /* ceil(n_bases/32) */
    uint32_t *base_2bits;      /* 2*n_words */
    uint32_t *base_3bits;      /* n_words */
    uint32_t *hash_count;      /* 4^k */
    uint32_t *hash_offset_end; /* 4^k + 1 (END offsets, as serial_index.c writes) */
    uint32_t *positions;       /* n_pos */
    uint16_t *pos_lseeds;      /* n_pos */
    uint64_t n_pos;
} swm_index_host_t;

/* Timing/counters of the last batch on a context. */
typedef struct swm_stats {
    uint64_t n_reads;
    uint64_t n_verifications;  /* (read, strand, candidate) windows run through banded Myers */
    uint64_t n_band_steps;     /* sum of read lengths over those verifications (Myers steps) */
    uint64_t n_records;
    uint64_t n_positions;      /* bucket entries fetched for the selected seeds (both strands) */
    /* stage times (HIP events on the context stream): seed lookup; candidate
     * batches; compaction; banded Myers; Hamming + traceback + emission */
    double ms_seed, ms_plan, ms_compact, ms_myers, ms_emit, ms_total;
    /* host wall clock of the call: enqueueing (input staging, copies,
     * launches, readbacks between stages) and the final wait */
    double ms_host_enqueue, ms_host_wait;
    /* with SWM_OPT_HISTOGRAM: reads by candidate count over both strands,
     * bin 0 = none, bin b = [2^(b-1), 2^b); and the largest count */
    uint64_t cand_hist[24];
    uint64_t max_cands;
    /* work-queue load: heavy (read, strand) items of the seed heap, batches
     * sorted by the large-batch kernel, reads emitted one wave each */
    uint64_t n_heavy_items, n_large_batches, n_heavy_reads;
    /* of those, the items and reads over the giant thresholds
     * (SWM_OPT_GIANT_POS / SWM_OPT_GIANT_CANDS), which are taken first */
    uint64_t n_giant_items, n_giant_reads;
    /* (read, strand) items whose selected 12-mers include a bucket that the
     * reference's index holds out of ascending order */
    uint64_t n_irregular_items;
    /* 1 when this run captured and instantiated its graph (graph form) */
    uint64_t n_graph_captures;
    /* bucket entries fetched by the light-item plan kernel (k_plan_pair), and
     * its own time (work-queue form, stage-timed runs; else 0) */
    uint64_t n_light_positions;
    double ms_plan_light;
} swm_stats_t;

typedef struct swm_ctx swm_ctx_t;

/* ---- index build (replaces serial_index.c main :818-865,
 *      construct_hash_index :359-445) ---- */

/* FASTA -> index files <prefix>.{ref_info,base_2bits,base_3bits,hash_count,
 * hash_offset,positions,pos_lseeds}, byte-identical to serial_index.c under
 * HOST_BUILD.  The heavy stages (encode, histogram, scan, stable scatter)
 * run on the GPU.  k is the -s option (2..14). */
int swm_index_build(const char *fasta_path, const char *prefix, uint32_t k);

/* In-memory variant: seq = contigs concatenated in FASTA order.  Fills *out
 * (free with swm_index_host_free).  The GPU stages report their time in *ms
 * when ms != NULL. */
int swm_index_build_mem(const char *seq, uint64_t n, uint32_t k, swm_index_host_t *out, double *ms);
void swm_index_host_free(swm_index_host_t *ix);

/* Write/read the reference's index files for an in-memory index.  ref_info
 * is written with one contig of the given name when contigs == NULL. */
int swm_index_write(const swm_index_host_t *ix, const char *prefix, const char *contig_name);

/* ---- FASTA/FASTQ records (replaces kseq_init / kseq_read / kseq_destroy,
 *      kseq.h:161-237, which the reference's mapper and index builder use) ---- */

typedef struct swm_seqfile swm_seqfile_t;
typedef struct swm_seqrec {
    char *name, *seq, *qual;            /* NUL-terminated; qual empty for FASTA */
    size_t name_len, seq_len, qual_len;
    size_t name_cap, seq_cap, qual_cap;
} swm_seqrec_t;

/* kseq's parse rules exactly: the name ends at the first isspace()
 * character, sequence and quality lines lose one trailing '\r', a header at
 * the end of the file is a record with an empty sequence.  r starts zeroed
 * and is reused record to record.  Returns the sequence length, -1 at the
 * end of the file, -2 for a quality string of another length, -3 when out
 * of memory (kseq_read's codes). */
swm_seqfile_t *swm_seqfile_open(const char *path);
int swm_seqfile_next(swm_seqfile_t *s, swm_seqrec_t *r);
void swm_seqfile_close(swm_seqfile_t *s);
void swm_seqrec_free(swm_seqrec_t *r);

/* ---- device-resident index (replaces load_ref_info, mapper_mpe.c:198-314) ---- */

/* Loads <prefix>.* with seed length k (-s) to the current HIP device. */
int swm_index_load(const char *prefix, uint32_t k, swm_ctx_t **ctx);
/* Same from memory: hash_count (4^k), positions/pos_lseeds (n_pos),
 * base_2bits (n_2bits_words as in the file). */
int swm_ctx_create(uint32_t k, const uint32_t *hash_count, const uint32_t *positions,
                   const uint16_t *pos_lseeds, uint64_t n_pos, const uint32_t *base_2bits,
                   uint64_t n_2bits_words, swm_ctx_t **ctx);
void swm_ctx_destroy(swm_ctx_t *ctx);

/* ---- mapping (replaces athread_spawn(find_positions)+athread_join,
 *      mapper_mpe.c:972-1036; kernel mapper_cpe.c:2592-2663) ---- */

/* Host buffers: reads[i*stride .. +lens[i]) for i < n (the reference's
 * reads[cur_block] / reads_seq_len[cur_block] with stride = -l), errors e
 * (-e, 0..15).  Requires lens[i] + 2 <= stride <= 256 (kseq stride rule).
 * out[n] receives one result per read. */
int swm_map_batch(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens,
                  uint64_t n, uint32_t e, swm_read_result_t *out);

/* Device buffers (HIP device pointers), enqueued on the context's stream;
 * returns when the results are in d_out.  For callers that keep reads
 * resident in HBM. */
int swm_map_batch_device(swm_ctx_t *ctx, const char *d_reads, uint64_t stride,
                         const uint32_t *d_lens, uint64_t n, uint32_t e,
                         swm_read_result_t *d_out);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  The same
 * mapping as swm_map_batch_device, with no allocation of any kind: its
 * intermediate buffers are carved from one static device array of
 * SWM_NOALLOC_SCRATCH_BYTES per GPU (held from the first context on that
 * GPU until the process ends).  n <= SWM_NOALLOC_MAX_READS; a batch whose
 * seed positions, bucket chunks or candidates exceed the scratch's share
 * returns SWM_ENOSPC with d_out unspecified (split it and call again).
 * Every call on one GPU uses that GPU's one array: calls that overlap on a
 * GPU (any contexts, any threads) race and must be serialised by the
 * caller.  Works in the form set by SWM_OPT_WORK_QUEUE; SWM_OPT_GRAPH and
 * SWM_OPT_HISTOGRAM are ignored. */
#define SWM_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define SWM_NOALLOC_MAX_READS 8192u
int swm_map_batch_device_noalloc(swm_ctx_t *ctx, const char *d_reads, uint64_t stride,
                                 const uint32_t *d_lens, uint64_t n, uint32_t e,
                                 swm_read_result_t *d_out);

/* Statistics of the last batch run on ctx. */
int swm_last_stats(const swm_ctx_t *ctx, swm_stats_t *st);

/* ---- alignment records and CIGAR text (the per-block output of put_cigar,
 *      mapper_cpe.c:1437-1507, and the CIGAR strings of align_reads,
 *      mapper_cpe.c:277-287, 475-507) ---- */

/* One written record.  The reference hands them to the host in blocks of at
 * most 256 per (read, strand) whose order depends on CPE timing; here they
 * come in one deterministic order: reads in batch order, the forward strand
 * before the reverse one, each strand's records in emission order. */
typedef struct swm_record {
    uint32_t read;       /* index of the read in the batch */
    uint32_t pos;        /* cigar->location: 1-based position in the concatenated reference */
    uint32_t cigar_off;  /* offset of the CIGAR text in swm_records_t.cigar */
    uint8_t strand;      /* 0 forward, 1 reverse */
    uint8_t dist;        /* edit distance (the CIGAR buffer's distance byte) */
    uint8_t primary;     /* 1 for the first record of its (read, strand): main_flag */
    uint8_t cigar_len;   /* bytes of CIGAR text (the CIGAR buffer's length byte) */
} swm_record_t;

typedef struct swm_records {
    uint64_t n;          /* records */
    swm_record_t *rec;
    uint64_t text_bytes;
    char *cigar;         /* the CIGAR texts back to back, no terminators */
    /* put_cigar's output blocks (mapper_cpe.c:1437-1507): the records of a
     * (read, strand) between two flushes (:1790, :1858, :2424) form one block;
     * block[b] is the index in rec of block b's first record (ascending).
     * swm_map_batch_records fills them; writers that only read rec and cigar
     * may leave them 0 / NULL. */
    uint64_t n_blocks;
    uint32_t *block;
} swm_records_t;

/* swm_map_batch that also returns every record with its CIGAR text (always
 * the synced form).  recs is filled with library-owned memory: release it
 * with swm_records_free. */
int swm_map_batch_records(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens,
                          uint64_t n, uint32_t e, swm_read_result_t *out, swm_records_t *recs);
void swm_records_free(swm_records_t *recs);

/* ---- the reference's own output buffers (for host code that keeps
 *      sam_format_2, mapper_mpe.c:426-687) ---- */

#define SWM_REF_MAX_CIGAR 256   /* MAX_CIGAR / CIGARS_PER_BLOCK (fm_define.h:17-18) */
#define SWM_REF_CIGAR_LEN 16    /* CIGAR_LEN / CIGAR_AVG_LEN (fm_define.h:19-20) */

/* One batch's ref_pos_buffer, cigar_buffer and cigar_num (the unmapped
 * list) as put_cigar (mapper_cpe.c:1437-1507) leaves them in the host
 * buffers of mapper_mpe.c:791-819, with cigar_block_size = n_blocks and
 * cigar_unmap_size = n_unmapped.  Block b: ref_pos[b * (SWM_REF_MAX_CIGAR +
 * 2)] = cur_read (2 * read + strand), [+1] = record count | main_flag << 31,
 * [+2 ..] = locations; cigar[b * SWM_REF_MAX_CIGAR * SWM_REF_CIGAR_LEN ..]
 * = per record its distance byte, text-length byte and CIGAR text.
 * unmapped[u] = 2 * read + 1 for each read with no record.  The reference
 * orders blocks and unmapped reads by CPE timing; here blocks follow the
 * records and unmapped reads the batch. */
typedef struct swm_ref_blocks {
    uint64_t n_blocks;
    uint32_t *ref_pos;
    char *cigar;
    uint64_t n_unmapped;
    uint32_t *unmapped;
} swm_ref_blocks_t;

/* From swm_map_batch_records' output (block starts included) for a batch
 * of n_reads reads.  SWM_EINVAL when the blocks do not tile the records
 * (one (read, strand) each, at most SWM_REF_MAX_CIGAR records) or a block's
 * text passes the reference's 4096 bytes per block.  Host only. */
int swm_records_to_ref_blocks(const swm_records_t *recs, uint64_t n_reads, swm_ref_blocks_t *out);
void swm_ref_blocks_free(swm_ref_blocks_t *b);

/* ---- SAM output (replaces sam_header and sam_format_2, mapper_mpe.c:315-687) ---- */

typedef struct swm_sam swm_sam_t;

/* Opens path for writing and writes the header (@HD, one @SQ per contig of
 * <index_prefix>.ref_info, @PG).  gzip_level 1..9 writes gzip (the
 * reference writes <out>.sam.gz at level 1); 0 writes plain text. */
int swm_sam_open(const char *path, const char *index_prefix, int gzip_level, swm_sam_t **out);
/* One batch: the unmapped reads (no record) in batch order, then every
 * record in recs' order.  names[i] (name_lens[i] bytes, or NUL-terminated
 * when name_lens is NULL), seqs/quals at i * stride with lens[i] bytes;
 * quals may be NULL (FASTA input: QUAL written as '*'). */
int swm_sam_write_batch(swm_sam_t *w, uint64_t n, const char *const *names, const uint32_t *name_lens,
                        const char *seqs, const char *quals, uint64_t stride, const uint32_t *lens,
                        const swm_records_t *recs);
/* Worker threads that format and compress a batch (default: the online
 * CPUs, at most 16).  The text is the same for any count. */
int swm_sam_set_threads(swm_sam_t *w, int n);
/* Flushes and closes; returns the first write error, if any. */
int swm_sam_close(swm_sam_t *w);

/* ---- forms and tunables (no counterpart in the reference: its one form is
 *      the CPE kernel with compile-time sizes, mapper_cpe.c:15-25) ---- */

#define SWM_OPT_WORK_QUEUE 1        /* 1 (default): heavy items go to device work queues; 0: plain form */
#define SWM_OPT_GRAPH 2             /* 1: fixed-capacity pipeline replayed as a hipGraph; 0 (default): synced */
#define SWM_OPT_HEAVY_CANDS 3       /* candidates above which an item is heavy (1..65535, default 64) */
#define SWM_OPT_MYERS_BLOCK 4       /* verification workgroup size: 64, 128 or 256 (default) */
#define SWM_OPT_PLAN_HEAVY_WAVES 5  /* persistent plan waves per CU (1..32, default 32) */
#define SWM_OPT_EMIT_HEAVY_WAVES 6  /* cooperative emission waves per CU (1..64, default 32) */
#define SWM_OPT_HISTOGRAM 7         /* 1: also collect the per-read candidate histogram (stats) */
#define SWM_OPT_MYERS_PACKED 8      /* 1 (default): for e <= 7, two candidates per lane in 16-bit halves */
#define SWM_OPT_GIANT_POS 9         /* heavy plan items with more positions go first (1..2^24, default 16384) */
#define SWM_OPT_GIANT_CANDS 10      /* heavy emission reads with more candidates go first (1..2^24, default 16384) */
#define SWM_OPT_SEED_GENERIC 11     /* 1: the seed kernel for any k, also where a fixed-k one exists (k = 12); default 0 */
#define SWM_OPT_PLAN_OVERLAP 12     /* 1: work-queue form runs the heavy-item plan kernels on a second stream
                                     * beside the light-item kernel; 0 (default): after it */
#define SWM_OPT_PLAN_LIGHT_WAVES 13 /* waves per CU of the light-item plan kernel (8..64, default 32) */
#define SWM_N_OPTIONS 13

/* Every form and tunable gives the same results; they change only speed. */
int swm_ctx_set_option(swm_ctx_t *ctx, int opt, int64_t value);
int swm_ctx_get_option(const swm_ctx_t *ctx, int opt, int64_t *value);

/* Autotuner: times the pipeline on the given sample (host buffers, as
 * swm_map_batch) for each candidate setting of the tunables, one at a time,
 * keeps the fastest on ctx and writes every option's value to chosen[10]
 * (indexed by SWM_OPT_* - 1) when chosen != NULL.  reps >= 1 timed runs per
 * setting (median). */
int swm_autotune(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens, uint64_t n,
                 uint32_t e, uint32_t reps, int64_t *chosen);

/* ---- serving loop: continuous batching over a resident index (grows the
 *      reference's double-buffered host pipeline, mapper_mpe.c:916-1106) ---- */

typedef struct swm_server swm_server_t;

typedef struct swm_server_config {
    uint32_t n_slots;      /* batches on the device at once, each on its own stream (1..16; 0 = 4) */
    uint32_t e;            /* -e */
    uint64_t stride;       /* -l: read stride of every submitted batch */
    uint64_t hbm_budget;   /* device bytes the admitted batches may hold (0 = 90% of free HBM at creation) */
    uint64_t max_batch;    /* reads per batch accepted (0 = 65536) */
    uint64_t max_launch;   /* continuous batching: a free slot takes every waiting batch that fits, up to
                            * this many reads, in one pipeline launch (0: one batch per launch) */
    uint32_t graph;        /* 1: replay captured graphs, launch sizes padded to powers of two (>= 1024) */
    uint32_t pad;
} swm_server_config_t;

typedef struct swm_batch_info {
    uint64_t ticket;
    uint64_t n_reads;
    double t_submit, t_admit, t_done;  /* seconds on CLOCK_MONOTONIC */
    int32_t status;                    /* SWM_OK or the mapping error */
    uint32_t slot;
    uint64_t est_bytes;                /* device bytes the batch was admitted against */
    uint64_t launch_reads;             /* reads of the launch the batch ran in (merged batches share one) */
    uint32_t launch_batches;           /* batches in that launch */
    uint32_t pad;
    swm_stats_t stats;                 /* of that launch */
} swm_batch_info_t;

/* Admission is first come first served: when a slot is free it takes the
 * oldest waiting batches -- as many as fit max_launch reads and the budget
 * (each batch held against its estimated device bytes) -- and maps them in
 * one launch.  Launches on different slots overlap on the device; each
 * result is the same as the batch mapped alone (swm_map_batch). */
int swm_server_create(swm_ctx_t *ctx, const swm_server_config_t *cfg, swm_server_t **out);
/* Non-blocking.  reads/lens/out stay owned by the caller and must stay valid
 * until swm_server_wait returns for the ticket.  SWM_EINVAL for a bad batch,
 * SWM_ENOMEM if the batch alone exceeds the budget. */
int swm_server_submit(swm_server_t *srv, const char *reads, const uint32_t *lens, uint64_t n,
                      swm_read_result_t *out, uint64_t *ticket);
/* Before traffic: maps the sample (repeated to size) on every slot at every
 * padded batch size up to max_batch, so that capacities are learned and
 * graphs captured up front.  Must not overlap submissions. */
int swm_server_warmup(swm_server_t *srv, const char *reads, const uint32_t *lens, uint64_t n);
/* Blocks until the batch is done; fills info (may be NULL) and forgets the ticket. */
int swm_server_wait(swm_server_t *srv, uint64_t ticket, swm_batch_info_t *info);
/* Totals since creation: batches, reads, peak admitted bytes, budget. */
int swm_server_totals(const swm_server_t *srv, uint64_t *batches, uint64_t *reads, uint64_t *peak_bytes,
                      uint64_t *budget);
/* Waits for every submitted batch, then stops the workers. */
void swm_server_destroy(swm_server_t *srv);

/* ---- stage entry points (per-kernel parity) ---- */

/* Banded Myers over explicit candidates, device pointers:
 * d_read_words[n_reads][16] interleaved 2-bit read words, d_lens[n_reads],
 * d_cand_read[n] (read row), d_win_start[n] (window start, genome bases),
 * d_ref_len[n]; outputs d_err[n], d_loc[n] (MAXERROR 10000 / -1 if > e). */
int swm_myers_device(swm_ctx_t *ctx, const uint32_t *d_read_words, const uint32_t *d_lens,
                     const uint32_t *d_cand_read, const uint32_t *d_win_start,
                     const uint32_t *d_ref_len, uint64_t n, uint32_t e,
                     int32_t *d_err, int32_t *d_loc);

/* Entry index of k-mer `code` (k = 2..14) in the device seed table, in
 * role 0 (keyed by its last k-1 bases) or 1 (by its first k-1 bases): the
 * table holds (offset, count) of every k-mer twice, so that two consecutive
 * k-mers of a read and the reverse complements of both share one 128-byte
 * line (16 entries of 8 bytes).  Host function, for layout checks; no
 * counterpart in the reference, whose hash_offset array
 * (mapper_mpe.c:250-255) is indexed by code. */
uint64_t swm_seed_slot(uint32_t code, uint32_t k, uint32_t role);

/* Measured int32 VALU throughput of the current device (lane ops/s), the
 * peak the Myers roofline is quoted against.  No counterpart in the
 * reference. */
int swm_valu_probe(double *ops_per_s);

/* Measured rate of random 8-byte gathers (one 128-byte line read each) over
 * a device array of array_bytes, in gathers per second: the ceiling the
 * gather-bound seed and bucket kernels are quoted against, beside the HBM
 * byte peak.  No counterpart in the reference. */
int swm_gather_probe(uint64_t array_bytes, double *gathers_per_s);

/* ---- device memory plumbing for callers without their own HIP setup
 *      (no counterpart in the reference, whose host/CPE memory is shared) ---- */
int swm_device_count(int *n);
int swm_dev_alloc(void **p, uint64_t bytes);
int swm_dev_free(void *p);
int swm_copy_h2d(void *dst, const void *src, uint64_t bytes);
int swm_copy_d2h(void *dst, const void *src, uint64_t bytes);
int swm_copy_d2d(void *dst, const void *src, uint64_t bytes);
int swm_dev_memset(void *dst, int value, uint64_t bytes);
/* page-locked host memory (DMA at full link speed; e.g. serving buffers) */
int swm_host_alloc(void **p, uint64_t bytes);
int swm_host_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
