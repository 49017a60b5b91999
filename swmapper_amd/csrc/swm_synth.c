/*
 * swm_synth.c -- seeded synthetic genomes and reads for the mapping workloads.
 *
 * Workload generator only (no counterpart in the reference, which ships no
 * data).  Deterministic for a given seed: xoshiro256** seeded by splitmix64.
 */
#include "swm_synth.h"

#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r)   /* xoshiro256** */
{
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

/* uniform in [0, n) (Lemire's multiply-shift; bias < 2^-32 for our n) */
static inline uint64_t rng_below(rng_t *r, uint64_t n) { return (uint64_t)(((__uint128_t)rng_next(r) * n) >> 64); }
static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static const char ACGT[4] = {'A', 'C', 'G', 'T'};

static inline char other_base(rng_t *r, char b)
{
    int idx = 0;
    switch (b) { case 'A': idx = 0; break; case 'C': idx = 1; break; case 'G': idx = 2; break; default: idx = 3; }
    return ACGT[(idx + 1 + (int)rng_below(r, 3)) & 3];
}

int swm_synth_genome(char *out, uint64_t n, uint64_t seed, int kind, double frac)
{
    if (!out || n == 0) return -1;
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = ACGT[rng_next(&r) >> 62];
    if (kind == SWM_GENOME_UNIFORM) return 0;
    uint64_t target = (uint64_t)(frac * (double)n);
    if (kind == SWM_GENOME_REPEAT_FAMILIES) {
        /* families of 300 bp units, 500 copies each, 10% per-base divergence */
        const uint32_t unit = 300, copies = 500;
        char cons[300];
        uint64_t placed = 0;
        while (placed < target && n > unit) {
            for (uint32_t j = 0; j < unit; j++) cons[j] = ACGT[rng_next(&r) >> 62];
            for (uint32_t c = 0; c < copies && placed < target; c++) {
                uint64_t at = rng_below(&r, n - unit);
                for (uint32_t j = 0; j < unit; j++)
                    out[at + j] = rng_unit(&r) < 0.10 ? other_base(&r, cons[j]) : cons[j];
                placed += unit;
            }
        }
        return 0;
    }
    if (kind == SWM_GENOME_TANDEM) {
        /* tandem arrays: period U[2,50], array length U[100, 6000], 2% divergence */
        char motif[64];
        uint64_t placed = 0;
        while (placed < target) {
            uint32_t period = 2 + (uint32_t)rng_below(&r, 49);
            uint64_t len = 100 + rng_below(&r, 5901);
            if (len >= n) len = n / 2;
            for (uint32_t j = 0; j < period; j++) motif[j] = ACGT[rng_next(&r) >> 62];
            uint64_t at = rng_below(&r, n - len);
            for (uint64_t j = 0; j < len; j++) {
                char b = motif[j % period];
                out[at + j] = rng_unit(&r) < 0.02 ? other_base(&r, b) : b;
            }
            placed += len;
        }
        return 0;
    }
    return -2;
}

static inline char complement(char b)
{
    switch (b) { case 'A': return 'T'; case 'C': return 'G'; case 'G': return 'C'; case 'T': return 'A'; default: return b; }
}

int swm_synth_reads(const char *genome, uint64_t n, uint64_t n_reads, uint32_t len_min, uint32_t len_max,
                    double err_min, double err_max, double indel_share, uint64_t seed,
                    char *reads, uint64_t stride, uint32_t *lens, uint32_t *true_pos, uint8_t *true_strand)
{
    if (!genome || !reads || !lens || len_min == 0 || len_min > len_max || len_max + 2 > stride) return -1;
    if (n < (uint64_t)len_max * 2 + 16) return -1;
    char buf[1024];
    for (uint64_t q = 0; q < n_reads; q++) {
        rng_t r;
        rng_seed(&r, seed ^ (0x5851f42d4c957f2dull * (q + 1)));
        uint32_t L = len_min + (uint32_t)rng_below(&r, (uint64_t)(len_max - len_min + 1));
        double err = err_min + (err_max - err_min) * rng_unit(&r);
        double p_sub = err * (1.0 - indel_share);
        double p_indel = err * indel_share;
        uint64_t start = rng_below(&r, n - 2 * (uint64_t)L - 8);
        uint64_t g = start;
        uint32_t m = 0;
        while (m < L && g < n) {
            double u = rng_unit(&r);
            if (u < p_indel) {
                uint32_t ilen = 1 + (uint32_t)rng_below(&r, 3);
                if (rng_next(&r) >> 63) {          /* insertion into the read */
                    for (uint32_t j = 0; j < ilen && m < L; j++) buf[m++] = ACGT[rng_next(&r) >> 62];
                } else {                           /* deletion from the read */
                    g += ilen;
                }
                continue;
            }
            char b = genome[g++];
            if (rng_unit(&r) < p_sub) b = other_base(&r, b);
            buf[m++] = b;
        }
        while (m < L) buf[m++] = ACGT[rng_next(&r) >> 62];
        int rc = (int)(rng_next(&r) >> 63);
        char *dst = reads + q * stride;
        if (rc) { for (uint32_t j = 0; j < L; j++) dst[j] = complement(buf[L - 1 - j]); }
        else memcpy(dst, buf, L);
        for (uint64_t j = L; j < stride; j++) dst[j] = 0;
        lens[q] = L;
        if (true_pos) true_pos[q] = (uint32_t)start;
        if (true_strand) true_strand[q] = (uint8_t)rc;
    }
    return 0;
}
