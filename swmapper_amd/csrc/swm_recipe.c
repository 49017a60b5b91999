/*
 * swm_recipe.c -- per-read-length recipe of the read encoding.
 *
 * get_seed_from_reads (mapper_cpe.c:1246-1364), driven seed by seed by
 * choose_seed (:1372-1379), writes its 16 interleaved 2-bit words through a
 * stateful accumulator whose stores depend only on the read length L and the
 * seed length k, never on the bases: a word may end up holding the plain 16
 * bases, nothing, or a few bases left in the accumulator (Q3).  Every base
 * always sits in its own slot (read_pos % 16).  The bit planes are only ever
 * OR-ed, so they hold exactly the encoded prefix of the read.
 *
 * This file executes that store sequence once per (L, k) over sets of base
 * indices instead of bits.  The result, for each word j, is the set of bases
 * it holds, expressed as three 32-bit lane masks over the plain packings of
 * words j, j+1 and j+2; the kernels apply it branch-free.
 */
#include <string.h>

#include "swm_recipe.h"

typedef struct { uint64_t b[4]; } bset_t;   /* bases 0..255 */

static void bs_clear(bset_t *s) { memset(s, 0, sizeof(*s)); }
static void bs_add(bset_t *s, uint32_t i) { s->b[i >> 6] |= 1ull << (i & 63); }
static int bs_has(const bset_t *s, uint32_t i) { return (int)((s->b[i >> 6] >> (i & 63)) & 1u); }

int swm_recipe_build(uint32_t L, uint32_t k, swm_recipe_t *out)
{
    memset(out, 0, sizeof(*out));
    if (L > SWM_RECIPE_MAX_L || k < 2 || k > 16) return -1;
    bset_t enc, w[16], planes;
    bs_clear(&enc);
    bs_clear(&planes);
    for (int j = 0; j < 16; j++) bs_clear(&w[j]);
    const uint64_t seed_len = k + 3;
    const uint64_t seeds_size = (L / seed_len) * 4;
    for (uint64_t sn = 0; sn < seeds_size; sn++) {
        uint64_t i = (sn % 4 == 0) ? 0 : k - 1;
        uint64_t roff = (sn / 4) * seed_len, soff = sn % 4;
        uint64_t rp = roff + soff + i;
        for (; i < k; i++) {
            rp = roff + soff + i;
            bs_add(&enc, (uint32_t)rp);
            bs_add(&planes, (uint32_t)rp);
            if (rp % 16 == 15) { w[rp / 16] = enc; bs_clear(&enc); }
        }
        if ((uint64_t)L - rp < k) {
            for (; rp < L; rp++) {
                bs_add(&enc, (uint32_t)rp);
                bs_add(&planes, (uint32_t)rp);
                if (rp % 16 == 15) { w[rp / 16] = enc; bs_clear(&enc); }
            }
            w[(rp - 1) / 16] = enc;
        }
    }
    /* planes: a prefix of the read */
    uint32_t n = 0;
    while (n < L && bs_has(&planes, n)) n++;
    for (uint32_t q = n; q < 256; q++)
        if (bs_has(&planes, q)) return -2;
    out->n_planes = n;
    for (int j = 0; j < 16; j++) {
        for (uint32_t q = 0; q < 256; q++) {
            if (!bs_has(&w[j], q)) continue;
            int32_t d = (int32_t)(q / 16) - j;
            if (d < 0 || d > 2) return -3;      /* outside the three-word window */
            out->mask[j][d] |= 3u << (2 * (q % 16));
        }
    }
    return 0;
}

void swm_recipe_apply(const swm_recipe_t *r, const uint32_t *plain16, uint32_t *words16)
{
    for (int j = 0; j < 16; j++) {
        uint32_t v = plain16[j] & r->mask[j][0];
        if (j + 1 < 16) v |= plain16[j + 1] & r->mask[j][1];
        if (j + 2 < 16) v |= plain16[j + 2] & r->mask[j][2];
        words16[j] = v;
    }
}

/* Test hook: the recipe applied on the host to one read (the kernels apply
 * the same table); lets the CPU suite check the recipe for every length. */
int swm_encode_read_host(const char *read, uint32_t L, uint32_t k, uint32_t *words32)
{
    swm_recipe_t r;
    if (swm_recipe_build(L, k, &r)) return -1;
    uint32_t plain[16];
    memset(plain, 0, sizeof(plain));
    memset(words32, 0, 32 * sizeof(uint32_t));
    for (uint32_t t = 0; t < L; t++) {
        uint32_t c;
        switch (read[t]) {
        case 'A': c = 0; break;
        case 'C': c = 1; break;
        case 'G': c = 2; break;
        case 'T': c = 3; break;
        default: c = 1; break;
        }
        plain[t / 16] |= c << (2 * (t % 16));
        if (t < r.n_planes) {
            words32[16 + (t >> 5) * 2] |= (c & 1u) << (t & 31);
            words32[17 + (t >> 5) * 2] |= (c >> 1) << (t & 31);
        }
    }
    swm_recipe_apply(&r, plain, words32);
    return 0;
}
