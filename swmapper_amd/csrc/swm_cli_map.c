/*
 * swmapper -- mapper CLI with the reference's options
 * (getopt "h::l::s::p:i::o::e:f::", mapper_mpe.c:717-760): -l<stride>
 * (max read length + 2, attached, default 120), -e <errors> (default 4),
 * -s<k> (default 12), -p <index prefix>, -i<reads FASTA/Q>, -o<out prefix>.
 * Batches of 10,000 reads (mapper_mpe.c:727) are mapped on the GPU; the
 * per-read results go to <out>.tsv (name, pos, strand, dist, count).
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
    snprintf(path, sizeof(path), "%s.tsv", outp);
    FILE *fo = fopen(path, "w");
    if (!fo) { fprintf(stderr, "swmapper: cannot open output\n"); swm_seqfile_close(sf); swm_ctx_destroy(ctx); return 1; }
    const uint64_t batch = 10000;
    char *reads = (char *)calloc(batch, (size_t)stride);
    uint32_t *lens = (uint32_t *)calloc(batch, 4);
    char **names = (char **)calloc(batch, sizeof(char *));
    swm_read_result_t *res = (swm_read_result_t *)calloc(batch, sizeof(swm_read_result_t));
    swm_seqrec_t rec;
    memset(&rec, 0, sizeof(rec));
    int l = 0, status = 0;
    while (l >= 0 && reads && lens && names && res) {
        uint64_t n = 0;
        while (n < batch && (l = swm_seqfile_next(sf, &rec)) >= 0) {
            if ((long)rec.seq_len + 2 > stride) { fprintf(stderr, "swmapper: read longer than -l - 2\n"); status = 1; l = -1; break; }
            memset(reads + n * stride, 0, (size_t)stride);
            memcpy(reads + n * stride, rec.seq, rec.seq_len);
            lens[n] = (uint32_t)rec.seq_len;
            free(names[n]);
            names[n] = strdup(rec.name);
            n++;
        }
        if (status) break;
        if (n == 0) break;
        rc = swm_map_batch(ctx, reads, (uint64_t)stride, lens, n, (uint32_t)e, res);
        if (rc != SWM_OK) { fprintf(stderr, "swmapper: map failed (status %d)\n", rc); status = 1; break; }
        for (uint64_t i = 0; i < n; i++) {
            if (res[i].count)
                fprintf(fo, "%s\t%u\t%c\t%u\t%u\n", names[i], res[i].pos, res[i].strand ? '-' : '+', res[i].dist, res[i].count);
            else
                fprintf(fo, "%s\t*\t*\t*\t0\n", names[i]);
        }
    }
    if (l < -1) status = 1;
    for (uint64_t i = 0; i < batch && names; i++) free(names[i]);
    free(names); free(reads); free(lens); free(res);
    swm_seqrec_free(&rec);
    fclose(fo);
    swm_seqfile_close(sf);
    swm_ctx_destroy(ctx);
    return status;
}
