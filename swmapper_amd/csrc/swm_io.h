/*
 * swm_io.h -- host-side FASTA/FASTQ reader and index-file helpers
 * (internal to libswm; the public ABI is include/swm.h).
 */
#ifndef SWM_IO_H
#define SWM_IO_H

#include <stddef.h>
#include <stdint.h>

#include "swm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct swm_seqfile swm_seqfile_t;
typedef struct swm_seqrec {
    char *name, *seq, *qual;
    size_t name_len, seq_len, qual_len;
    size_t name_cap, seq_cap, qual_cap;
} swm_seqrec_t;

swm_seqfile_t *swm_seqfile_open(const char *path);
/* >= 0 sequence length, -1 end of file, -2 truncated quality, -3 no memory */
int swm_seqfile_next(swm_seqfile_t *s, swm_seqrec_t *r);
void swm_seqfile_close(swm_seqfile_t *s);
void swm_seqrec_free(swm_seqrec_t *r);

int swm_index_write_contigs(const swm_index_host_t *ix, const char *prefix, uint32_t n_contigs,
                            const uint64_t *lens, const char *const *names);

#ifdef __cplusplus
}
#endif
#endif
