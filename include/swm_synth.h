/*
 * swm_synth.h -- seeded synthetic genomes and reads (workload generator).
 *
 * No counterpart in the reference (Xu-Kai/SWMapper ships no data); these
 * produce the inputs of the mapping workloads with fixed seeds.
 */
#ifndef SWM_SYNTH_H
#define SWM_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SWM_GENOME_UNIFORM = 0,          /* i.i.d. uniform ACGT */
    SWM_GENOME_REPEAT_FAMILIES = 1,  /* + frac of 300 bp units x 500 copies at 10% divergence */
    SWM_GENOME_TANDEM = 2            /* + frac of tandem arrays, period U[2,50], 2% divergence */
};

/* Fill out[0..n) with ASCII ACGT.  Returns 0 on success. */
int swm_synth_genome(char *out, uint64_t n, uint64_t seed, int kind, double frac);

/* Sample n_reads reads uniformly from both strands.  Per read: length
 * U[len_min, len_max], error rate U[err_min, err_max] split into
 * substitutions (1 - indel_share) and indel events (indel_share) of length
 * U{1,2,3}, insertions and deletions 50/50.  reads: n_reads * stride bytes,
 * zero padded; lens[n_reads]; true_pos/true_strand optional. */
int swm_synth_reads(const char *genome, uint64_t n, uint64_t n_reads, uint32_t len_min, uint32_t len_max,
                    double err_min, double err_max, double indel_share, uint64_t seed,
                    char *reads, uint64_t stride, uint32_t *lens, uint32_t *true_pos, uint8_t *true_strand);

#ifdef __cplusplus
}
#endif
#endif
