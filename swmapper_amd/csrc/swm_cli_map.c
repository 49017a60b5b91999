/*
 * swmapper -- mapper CLI with the reference's options
 * (getopt "h::l::s::p:i::o::e:f::", mapper_mpe.c:717-760): -l<stride>
 * (max read length + 2, attached, default 120), -e <errors> (default 4),
 * -s<k> (default 12), -p <index prefix>, -i<reads FASTA/Q>, -o<out prefix>.
 * Batches of 10,000 reads (mapper_mpe.c:727) are mapped on the GPU and
 * written as SAM to <out>.sam.gz (gzip level 1), as the reference writes
 * them (mapper_mpe.c:847, 866): header, then per batch the unmapped reads
 * and every record with its CIGAR.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "swm.h"
#include "swm_io.h"

int main(int argc, char **argv)
{
    long stride = 120, e = 4, k = 12;
    const char *prefix = NULL, *in = NULL, *outp = NULL;
    int opt;
    while ((opt = getopt(argc, argv, "h::l::s::p:i::o::e:f::")) != -1) {
        switch (opt) {
        case 'l': if (optarg) stride = strtol(optarg, NULL, 10); break;
        case 'p': prefix = optarg; break;
        case 's': if (optarg) k = strtol(optarg, NULL, 10); break;
        case 'o': outp = optarg; break;
        case 'e': e = strtol(optarg, NULL, 10); break;
        case 'i': in = optarg; break;
        case 'h':
            printf("-l\tmax read length\n-o\toutput file prefix\n-e\terror size\n-p\tindex prefix\n"
                   "-s\t[opt, default:12] seed length\n-i\tsingle data input\n");
            return 0;
        default: return 2;
        }
    }
    if (!prefix || !in || !outp || stride < 3 || stride > 256 || e < 0 || e > 15 || k < 2 || k > 14) {
        fprintf(stderr, "swmapper: need -p -i -o with -l in [3,256], -e in [0,15]\n");
        return 2;
    }
    swm_ctx_t *ctx = NULL;
    int rc = swm_index_load(prefix, (uint32_t)k, &ctx);
    if (rc != SWM_OK) { fprintf(stderr, "swmapper: index load failed (status %d)\n", rc); return 1; }
    swm_seqfile_t *sf = swm_seqfile_open(in);
    if (!sf) { fprintf(stderr, "swmapper: cannot open reads\n"); swm_ctx_destroy(ctx); return 1; }
    char path[4096];
    snprintf(path, sizeof(path), "%s.sam.gz", outp);
    swm_sam_t *sam = NULL;
    rc = swm_sam_open(path, prefix, 1, &sam);
    if (rc != SWM_OK) {
        fprintf(stderr, "swmapper: cannot open output (status %d)\n", rc);
        swm_seqfile_close(sf);
        swm_ctx_destroy(ctx);
        return 1;
    }
    const uint64_t batch = 10000;
    char *reads = (char *)calloc(batch, (size_t)stride);
    char *quals = (char *)calloc(batch, (size_t)stride);
    uint32_t *lens = (uint32_t *)calloc(batch, 4);
    char **names = (char **)calloc(batch, sizeof(char *));
    swm_read_result_t *res = (swm_read_result_t *)calloc(batch, sizeof(swm_read_result_t));
    swm_seqrec_t rec;
    memset(&rec, 0, sizeof(rec));
    int l = 0, status = 0;
    while (l >= 0 && reads && quals && lens && names && res) {
        uint64_t n = 0;
        int all_qual = 1;
        while (n < batch && (l = swm_seqfile_next(sf, &rec)) >= 0) {
            if ((long)rec.seq_len + 2 > stride) { fprintf(stderr, "swmapper: read longer than -l - 2\n"); status = 1; l = -1; break; }
            memset(reads + n * stride, 0, (size_t)stride);
            memcpy(reads + n * stride, rec.seq, rec.seq_len);
            memset(quals + n * stride, 0, (size_t)stride);
            if (rec.qual_len == rec.seq_len && rec.qual) memcpy(quals + n * stride, rec.qual, rec.qual_len);
            else all_qual = 0;
            lens[n] = (uint32_t)rec.seq_len;
            free(names[n]);
            names[n] = strdup(rec.name);
            n++;
        }
        if (status) break;
        if (n == 0) break;
        swm_records_t recs;
        rc = swm_map_batch_records(ctx, reads, (uint64_t)stride, lens, n, (uint32_t)e, res, &recs);
        if (rc != SWM_OK) { fprintf(stderr, "swmapper: map failed (status %d)\n", rc); status = 1; break; }
        rc = swm_sam_write_batch(sam, n, (const char *const *)names, NULL, reads, all_qual ? quals : NULL,
                                 (uint64_t)stride, lens, &recs);
        swm_records_free(&recs);
        if (rc != SWM_OK) { fprintf(stderr, "swmapper: write failed (status %d)\n", rc); status = 1; break; }
    }
    if (l < -1) status = 1;
    for (uint64_t i = 0; i < batch && names; i++) free(names[i]);
    free(names); free(reads); free(quals); free(lens); free(res);
    swm_seqrec_free(&rec);
    if (swm_sam_close(sam) != SWM_OK) status = 1;
    swm_seqfile_close(sf);
    swm_ctx_destroy(ctx);
    return status;
}
