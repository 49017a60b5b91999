/*
 * swm_io.h -- host-side index-file helpers (internal to libswm; the
 * public ABI, the FASTA/FASTQ reader included, is include/swm.h).
 */
#ifndef SWM_IO_H
#define SWM_IO_H

#include <stddef.h>
#include <stdint.h>

#include "swm.h"

#ifdef __cplusplus
extern "C" {
#endif

int swm_index_write_contigs(const swm_index_host_t *ix, const char *prefix, uint32_t n_contigs,
                            const uint64_t *lens, const char *const *names);

#ifdef __cplusplus
}
#endif
#endif
