/*
 * swm_recipe.h -- per-read-length recipe of the read encoding (internal).
 */
#ifndef SWM_RECIPE_H
#define SWM_RECIPE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWM_RECIPE_MAX_L 254

typedef struct swm_recipe {
    uint32_t n_planes;      /* bases 0..n_planes-1 are in the bit planes */
    uint32_t mask[16][3];   /* word j = plain[j]&mask[j][0] | plain[j+1]&mask[j][1] | plain[j+2]&mask[j][2] */
} swm_recipe_t;

/* 0 on success; negative if the store pattern leaves the assumed shape. */
int swm_recipe_build(uint32_t L, uint32_t k, swm_recipe_t *out);
/* plain16: plain 2-bit packing of the read (16 bases per word, zeros past L). */
void swm_recipe_apply(const swm_recipe_t *r, const uint32_t *plain16, uint32_t *words16);

/* Exported for the CPU tests: encode one read with the recipe. */
int swm_encode_read_host(const char *read, uint32_t L, uint32_t k, uint32_t *words32);

#ifdef __cplusplus
}
#endif
#endif
