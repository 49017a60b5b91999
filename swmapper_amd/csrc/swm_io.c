/*
 * swm_io.c -- host side of the boundary: FASTA/FASTQ parsing and the
 * reference's index files.
 *
 * File layout (all little-endian, headerless), written as serial_index.c
 * writes it and read as load_ref_info reads it:
 *   <p>.ref_info    text, per contig: length, name length, name
 *                   (serial_index.c:406-408; mapper_mpe.c:213-225)
 *   <p>.base_2bits  u32 x 2*ceil(N/32)   (serial_index.c:73-81)
 *   <p>.base_3bits  u32 x ceil(N/32)     (:131-135)
 *   <p>.hash_count  u32 x 4^k            (:98-102)
 *   <p>.hash_offset u32 x 4^k+1, END offsets after the scatter (:105-109)
 *   <p>.positions   u32 x M              (:84-90)
 *   <p>.pos_lseeds  u16 x M              (:117-121)
 */
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "swm.h"
#include "swm_io.h"

/* ---------------------------------------------------------------- FASTA/Q */

/* Record reader with kseq's rules (kseq_read, kseq.h:197-237, over
 * ks_getuntil2, :99-147):
 *   - records start at '>' or '@' (anything before the first is skipped);
 *   - the name runs to the first isspace() character (' ', \t, \n, \v, \f,
 *     \r); unless that was '\n' the rest of the line is the comment;
 *   - a sequence line starting with '>', '@' or '+' ends the sequence,
 *     empty lines are skipped, other lines are appended whole; after each
 *     line one trailing '\r' is dropped when the sequence so far is longer
 *     than one character (so a CR-only first line keeps its '\r');
 *   - after '+' the rest of that line is skipped and whole quality lines
 *     are appended (with the same '\r' rule) until there are at least as
 *     many as sequence characters; a different count is an error (-2);
 *   - a header with nothing after it at the end of the file is a record
 *     with an empty sequence. */
struct swm_seqfile {
    FILE *f;
    int last;    /* first char of the next header, already consumed */
};

swm_seqfile_t *swm_seqfile_open(const char *path)
{
    if (!path) return NULL;
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    swm_seqfile_t *s = (swm_seqfile_t *)calloc(1, sizeof(*s));
    if (!s) { fclose(f); return NULL; }
    s->f = f;
    return s;
}

void swm_seqfile_close(swm_seqfile_t *s)
{
    if (!s) return;
    fclose(s->f);
    free(s);
}

static int grow(char **buf, size_t *cap, size_t need)
{
    if (need <= *cap) return 0;
    size_t nc = *cap ? *cap : 256;
    while (nc < need) nc *= 2;
    char *nb = (char *)realloc(*buf, nc);
    if (!nb) return -1;
    *buf = nb;
    *cap = nc;
    return 0;
}

/* ks_getuntil2(ks, KS_SEP_LINE, str, 0, append = 1): the rest of the line
 * appended; -1 when the file ended before any character, -3 on no memory. */
static int line_append(FILE *f, char **buf, size_t *len, size_t *cap)
{
    int c, got = 0;
    while ((c = fgetc(f)) != EOF) {
        got = 1;
        if (c == '\n') break;
        if (grow(buf, cap, *len + 2)) return -3;
        (*buf)[(*len)++] = (char)c;
    }
    if (!got) return -1;
    if (*len > 1 && (*buf)[*len - 1] == '\r') --*len;
    return 0;
}

int swm_seqfile_next(swm_seqfile_t *s, swm_seqrec_t *r)
{
    if (!s || !r) return -1;
    int c = 0;
    if (s->last == 0) {
        while ((c = fgetc(s->f)) != EOF && c != '>' && c != '@') {}
        if (c == EOF) return -1;
        s->last = c;
    }
    r->name_len = r->seq_len = r->qual_len = 0;
    if (grow(&r->name, &r->name_cap, 2) || grow(&r->seq, &r->seq_cap, 2) || grow(&r->qual, &r->qual_cap, 2))
        return -3;
    /* the name: ks_getuntil(ks, KS_SEP_SPACE, ...) */
    int got = 0, delim = 0;
    while ((c = fgetc(s->f)) != EOF) {
        got = 1;
        if (isspace(c)) { delim = c; break; }
        if (grow(&r->name, &r->name_cap, r->name_len + 2)) return -3;
        r->name[r->name_len++] = (char)c;
    }
    r->name[r->name_len] = 0;
    if (!got) return -1;
    if (delim != '\n') while ((c = fgetc(s->f)) != EOF && c != '\n') {}   /* the comment */
    while ((c = fgetc(s->f)) != EOF && c != '>' && c != '+' && c != '@') {
        if (c == '\n') continue;
        if (grow(&r->seq, &r->seq_cap, r->seq_len + 2)) return -3;
        r->seq[r->seq_len++] = (char)c;
        if (line_append(s->f, &r->seq, &r->seq_len, &r->seq_cap) == -3) return -3;
    }
    s->last = (c == '>' || c == '@') ? c : 0;
    r->seq[r->seq_len] = 0;
    r->qual[0] = 0;
    if (c != '+') return (int)(r->seq_len > 0x7fffffff ? 0x7fffffff : r->seq_len);
    while ((c = fgetc(s->f)) != EOF && c != '\n') {}
    if (c == EOF) return -2;
    int q;
    while ((q = line_append(s->f, &r->qual, &r->qual_len, &r->qual_cap)) >= 0 && r->qual_len < r->seq_len) {}
    if (q == -3) return -3;
    r->qual[r->qual_len] = 0;
    s->last = 0;
    if (r->qual_len != r->seq_len) return -2;
    return (int)(r->seq_len > 0x7fffffff ? 0x7fffffff : r->seq_len);
}

void swm_seqrec_free(swm_seqrec_t *r)
{
    free(r->name); free(r->seq); free(r->qual);
    memset(r, 0, sizeof(*r));
}

/* ------------------------------------------------------------ index files */

static int write_file(const char *prefix, const char *ext, const void *p, size_t elem, size_t n)
{
    char path[4096];
    if (snprintf(path, sizeof(path), "%s%s", prefix, ext) >= (int)sizeof(path)) return SWM_EINVAL;
    FILE *f = fopen(path, "wb");
    if (!f) return SWM_EIO;
    size_t w = n ? fwrite(p, elem, n, f) : 0;
    int rc = (w == n) ? SWM_OK : SWM_EIO;
    if (fclose(f) != 0) rc = SWM_EIO;
    return rc;
}

int swm_index_write_contigs(const swm_index_host_t *ix, const char *prefix, uint32_t n_contigs,
                            const uint64_t *lens, const char *const *names)
{
    if (!ix || !prefix) return SWM_EINVAL;
    char path[4096];
    if (snprintf(path, sizeof(path), "%s.ref_info", prefix) >= (int)sizeof(path)) return SWM_EINVAL;
    FILE *f = fopen(path, "w");
    if (!f) return SWM_EIO;
    for (uint32_t i = 0; i < n_contigs; i++)
        fprintf(f, "%lu\n%lu\n%s\n", (unsigned long)lens[i], (unsigned long)strlen(names[i]), names[i]);
    if (fclose(f) != 0) return SWM_EIO;
    const uint64_t nb = 1ull << (2 * ix->k);
    int rc;
    if ((rc = write_file(prefix, ".base_2bits", ix->base_2bits, 4, 2 * ix->n_words))) return rc;
    if ((rc = write_file(prefix, ".positions", ix->positions, 4, ix->n_pos))) return rc;
    if ((rc = write_file(prefix, ".hash_count", ix->hash_count, 4, nb))) return rc;
    if ((rc = write_file(prefix, ".hash_offset", ix->hash_offset_end, 4, nb + 1))) return rc;
    if ((rc = write_file(prefix, ".pos_lseeds", ix->pos_lseeds, 2, ix->n_pos))) return rc;
    if ((rc = write_file(prefix, ".base_3bits", ix->base_3bits, 4, ix->n_words))) return rc;
    return SWM_OK;
}

int swm_index_write(const swm_index_host_t *ix, const char *prefix, const char *contig_name)
{
    if (!ix) return SWM_EINVAL;
    const char *nm = contig_name ? contig_name : "chr1";
    uint64_t len = ix->n_bases;
    return swm_index_write_contigs(ix, prefix, 1, &len, &nm);
}

static void *read_file(const char *prefix, const char *ext, size_t *bytes)
{
    char path[4096];
    if (snprintf(path, sizeof(path), "%s%s", prefix, ext) >= (int)sizeof(path)) return NULL;
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    if (fseek(f, 0, SEEK_END) != 0) { fclose(f); return NULL; }
    long sz = ftell(f);
    if (sz < 0 || fseek(f, 0, SEEK_SET) != 0) { fclose(f); return NULL; }
    void *p = malloc((size_t)sz + 64);
    if (!p) { fclose(f); return NULL; }
    if (sz && fread(p, 1, (size_t)sz, f) != (size_t)sz) { free(p); fclose(f); return NULL; }
    fclose(f);
    memset((char *)p + sz, 0, 64);
    *bytes = (size_t)sz;
    return p;
}

int swm_index_load(const char *prefix, uint32_t k, swm_ctx_t **ctx)
{
    if (!prefix || !ctx || k < 2 || k > 14) return SWM_EINVAL;
    size_t b_cnt = 0, b_pos = 0, b_ls = 0, b_b2 = 0;
    uint32_t *cnt = (uint32_t *)read_file(prefix, ".hash_count", &b_cnt);
    uint32_t *pos = (uint32_t *)read_file(prefix, ".positions", &b_pos);
    uint16_t *ls = (uint16_t *)read_file(prefix, ".pos_lseeds", &b_ls);
    uint32_t *b2 = (uint32_t *)read_file(prefix, ".base_2bits", &b_b2);
    int rc = SWM_OK;
    const uint64_t nb = 1ull << (2 * k);
    if (!cnt || !pos || !ls || !b2) rc = SWM_EIO;
    /* mis-sized files are rejected: the reference would read past them */
    else if (b_cnt != nb * 4 || b_pos % 4 || b_ls % 2 || b_pos / 4 != b_ls / 2 || b_b2 % 8)
        rc = SWM_EIO;
    else {
        uint64_t tot = 0;
        for (uint64_t i = 0; i < nb; i++) tot += cnt[i];
        if (tot != b_pos / 4) rc = SWM_EIO;
    }
    if (rc == SWM_OK) rc = swm_ctx_create(k, cnt, pos, ls, b_pos / 4, b2, b_b2 / 4, ctx);
    free(cnt); free(pos); free(ls); free(b2);
    return rc;
}

int swm_index_build(const char *fasta_path, const char *prefix, uint32_t k)
{
    if (!fasta_path || !prefix || k < 2 || k > 14) return SWM_EINVAL;
    swm_seqfile_t *sf = swm_seqfile_open(fasta_path);
    if (!sf) return SWM_EIO;
    swm_seqrec_t rec;
    memset(&rec, 0, sizeof(rec));
    char *seq = NULL;
    size_t n = 0, cap = 0;
    uint64_t *lens = NULL;
    char **names = NULL;
    uint32_t nc = 0;
    int rc = SWM_OK, l;
    while ((l = swm_seqfile_next(sf, &rec)) >= 0) {
        if (grow(&seq, &cap, n + rec.seq_len + 1)) { rc = SWM_ENOMEM; break; }
        memcpy(seq + n, rec.seq, rec.seq_len);
        n += rec.seq_len;
        uint64_t *nl = (uint64_t *)realloc(lens, (nc + 1) * sizeof(uint64_t));
        char **nn = (char **)realloc(names, (nc + 1) * sizeof(char *));
        if (!nl || !nn) { free(nl ? nl : lens); free(nn ? nn : names); lens = NULL; names = NULL; rc = SWM_ENOMEM; break; }
        lens = nl; names = nn;
        lens[nc] = rec.seq_len;
        names[nc] = strdup(rec.name);
        nc++;
    }
    if (rc == SWM_OK && l < -1) rc = SWM_EIO;
    swm_seqfile_close(sf);
    swm_seqrec_free(&rec);
    if (rc == SWM_OK && (nc == 0 || n < k)) rc = SWM_EIO;
    swm_index_host_t ix;
    memset(&ix, 0, sizeof(ix));
    if (rc == SWM_OK) rc = swm_index_build_mem(seq, n, k, &ix, NULL);
    if (rc == SWM_OK) rc = swm_index_write_contigs(&ix, prefix, nc, lens, (const char *const *)names);
    swm_index_host_free(&ix);
    for (uint32_t i = 0; i < nc; i++) free(names[i]);
    free(names); free(lens); free(seq);
    return rc;
}
