/*
 * swm_sam.c -- SAM output of mapped batches: the header of sam_header
 * (mapper_mpe.c:315-341) and the lines of sam_format_2 (:426-687), written
 * from the records of swm_map_batch_records.
 *
 * Line layout (fields separated by '\t', SEPARATOR :66):
 *   unmapped (:446-482):  NAME 4 * 0 0 * * 0 0 SEQ QUAL
 *   record   (:499-686):  NAME FLAG RNAME POS 255 CIGAR * 0 0 SEQ QUAL NM:i:d XM:i:d
 * FLAG is 0 / 256 on the forward strand and 16 / 272 on the reverse one
 * (primary / not); SEQ and QUAL are given for the primary record only (the
 * reverse complement and reversed qualities on the reverse strand, :342-366),
 * '*' otherwise.  RNAME / POS come from the record's 1-based location loc:
 * the contig c with start[c] <= loc < start[c + 1] over the cumulative
 * contig starts (get_ref_name's search, :367-397, always lands there), and
 * POS = loc - start[c].  So the last base of a contig is reported as base 0
 * of the next one, and the last base of the reference as base 0 of an
 * unnamed contig, as the reference does.  The distance digit is '0' + d
 * (a single character, also for d >= 10).
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <zlib.h>

#include "swm.h"

struct swm_sam {
    FILE *f;
    gzFile gz;
    int err;
    uint32_t n_contigs;
    uint32_t *start;        /* n_contigs + 1 cumulative starts (u32, as the reference's ref_len) */
    char **name;            /* n_contigs + 1 (the last one empty) */
    uint32_t *name_len;     /* bytes written for RNAME (the .ref_info name-length line) */
    char *line;
    size_t cap;
    char *rseq, *rqual;     /* reverse-complement scratch */
    size_t rcap;
};

static void out(swm_sam_t *w, const char *p, size_t n)
{
    if (w->err || n == 0) return;
    if (w->gz) {
        if (gzwrite(w->gz, p, (unsigned)n) != (int)n) w->err = SWM_EIO;
    } else if (fwrite(p, 1, n, w->f) != n) {
        w->err = SWM_EIO;
    }
}

/* One line buffer, grown to fit the longest line. */
static char *need(swm_sam_t *w, size_t n)
{
    if (n > w->cap) {
        size_t c = w->cap ? w->cap : 1024;
        while (c < n) c *= 2;
        char *p = (char *)realloc(w->line, c);
        if (!p) return NULL;
        w->line = p;
        w->cap = c;
    }
    return w->line;
}

/* append_int (mapper_mpe.c:410-425) on an int: digits of n, most
 * significant first; "0" for 0.  C's truncating division is kept, so a
 * negative n gives the same characters the reference prints. */
static char *put_int(char *o, int n)
{
    if (n == 0) { *o++ = '0'; return o; }
    char t[20];
    int i = 19;
    while (n) {
        t[i--] = (char)('0' + n % 10);
        n /= 10;
    }
    memcpy(o, t + i + 1, (size_t)(19 - i));
    return o + (19 - i);
}

static char *put(char *o, const char *s, size_t n)
{
    memcpy(o, s, n);
    return o + n;
}

#define PUTS(o, lit) put((o), (lit), sizeof(lit) - 1)

static int load_contigs(swm_sam_t *w, const char *prefix)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s.ref_info", prefix);
    FILE *f = fopen(path, "r");
    if (!f) return SWM_EIO;
    char *l = NULL;
    size_t cap = 0;
    uint32_t n = 0, m = 0;
    int rc = SWM_OK;
    for (;;) {
        ssize_t a = getline(&l, &cap, f);
        if (a < 0) break;
        uint32_t len = (uint32_t)strtoul(l, NULL, 10);
        ssize_t b = getline(&l, &cap, f);
        if (b < 0) { rc = SWM_EIO; break; }
        uint32_t nl = (uint32_t)strtoul(l, NULL, 10);
        ssize_t c = getline(&l, &cap, f);
        if (c < 0) { rc = SWM_EIO; break; }
        if (c > 0 && l[c - 1] == '\n') l[--c] = 0;
        if (n + 2 > m) {
            m = m ? 2 * m : 16;
            uint32_t *s2 = (uint32_t *)realloc(w->start, m * sizeof(uint32_t));
            if (s2) w->start = s2;
            char **n2 = (char **)realloc(w->name, m * sizeof(char *));
            if (n2) w->name = n2;
            uint32_t *l2 = (uint32_t *)realloc(w->name_len, m * sizeof(uint32_t));
            if (l2) w->name_len = l2;
            if (!s2 || !n2 || !l2) { rc = SWM_ENOMEM; break; }
            if (n == 0) w->start[0] = 0;
        }
        /* the name buffer holds at least name_len bytes, zero past the name */
        size_t nb = (size_t)c > nl ? (size_t)c : nl;
        w->name[n] = (char *)calloc(nb + 1, 1);
        if (!w->name[n]) { rc = SWM_ENOMEM; break; }
        memcpy(w->name[n], l, (size_t)c);
        w->name_len[n] = nl;
        w->start[n + 1] = w->start[n] + len;
        n++;
        w->n_contigs = n;
    }
    free(l);
    fclose(f);
    if (rc == SWM_OK && n == 0) rc = SWM_EINVAL;
    if (rc == SWM_OK) {
        w->name[n] = (char *)calloc(1, 1);
        w->name_len[n] = 0;
        if (!w->name[n]) rc = SWM_ENOMEM;
    }
    return rc;
}

static void free_writer(swm_sam_t *w)
{
    if (w->name)
        for (uint32_t i = 0; i <= w->n_contigs; i++) free(w->name[i]);
    free(w->name);
    free(w->name_len);
    free(w->start);
    free(w->line);
    free(w->rseq);
    free(w->rqual);
    free(w);
}

int swm_sam_open(const char *path, const char *index_prefix, int gzip_level, swm_sam_t **outp)
{
    if (!path || !index_prefix || !outp || gzip_level < 0 || gzip_level > 9) return SWM_EINVAL;
    *outp = NULL;
    swm_sam_t *w = (swm_sam_t *)calloc(1, sizeof(*w));
    if (!w) return SWM_ENOMEM;
    int rc = load_contigs(w, index_prefix);
    if (rc != SWM_OK) { free_writer(w); return rc; }
    if (gzip_level > 0) {
        char mode[4] = {'w', (char)('0' + gzip_level), 0, 0};
        w->gz = gzopen(path, mode);
        if (!w->gz) { free_writer(w); return SWM_EIO; }
    } else {
        w->f = fopen(path, "wb");
        if (!w->f) { free_writer(w); return SWM_EIO; }
    }
    /* header (mapper_mpe.c:315-341) */
    char *o = need(w, 256);
    if (!o) { swm_sam_close(w); return SWM_ENOMEM; }
    static const char hd[] = "@HD\tVN:1.4\tSO:unsorted\n";
    out(w, hd, sizeof(hd) - 1);
    for (uint32_t i = 0; i < w->n_contigs; i++) {
        const size_t nl = strlen(w->name[i]);
        o = need(w, nl + 64);
        if (!o) { swm_sam_close(w); return SWM_ENOMEM; }
        char *p = PUTS(o, "@SQ\tSN:");
        p = put(p, w->name[i], nl);
        p += sprintf(p, "\tLN:%u\n", (unsigned)(w->start[i + 1] - w->start[i]));
        out(w, o, (size_t)(p - o));
    }
    static const char pg[] = "@PG\tID:XXX\tVN:1.0\tPN:XXX\tCL:\"CMDLINE\"\n";
    out(w, pg, sizeof(pg) - 1);
    if (w->err) { rc = w->err; swm_sam_close(w); return rc; }
    *outp = w;
    return SWM_OK;
}

/* The contig of a 1-based location: the last c with start[c] <= loc. */
static uint32_t contig_of(const swm_sam_t *w, uint32_t loc)
{
    uint32_t lo = 0, hi = w->n_contigs + 1;       /* start[0..n_contigs] */
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (w->start[mid] <= loc) lo = mid;
        else hi = mid;
    }
    return lo;
}

static uint32_t name_bytes(const char *const *names, const uint32_t *name_lens, uint64_t i)
{
    return name_lens ? name_lens[i] : (uint32_t)strlen(names[i]);
}

int swm_sam_write_batch(swm_sam_t *w, uint64_t n, const char *const *names, const uint32_t *name_lens,
                        const char *seqs, const char *quals, uint64_t stride, const uint32_t *lens,
                        const swm_records_t *recs)
{
    if (!w || (n && (!names || !seqs || !lens)) || !recs || (recs->n && (!recs->rec || !recs->cigar)))
        return SWM_EINVAL;
    if (w->err) return w->err;
    for (uint64_t r = 0; r < recs->n; r++) {
        const swm_record_t *x = &recs->rec[r];
        if (x->read >= n || (uint64_t)x->cigar_off + x->cigar_len > recs->text_bytes) return SWM_EINVAL;
        if (r && (x->read < recs->rec[r - 1].read)) return SWM_EINVAL;   /* records come in read order */
    }
    /* unmapped reads: no record (put_cigar's unmap case, cigar_size == 0) */
    uint64_t r = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (r < recs->n && recs->rec[r].read == i) {
            while (r < recs->n && recs->rec[r].read == i) r++;
            continue;
        }
        const uint32_t nl = name_bytes(names, name_lens, i), L = lens[i];
        char *o = need(w, (size_t)nl + 2 * (size_t)L + 64);
        if (!o) return SWM_ENOMEM;
        char *p = put(o, names[i], nl);
        p = PUTS(p, "\t4\t*\t0\t0\t*\t*\t0\t0\t");
        p = put(p, seqs + i * stride, L);
        *p++ = '\t';
        if (quals) p = put(p, quals + i * stride, L);
        else *p++ = '*';
        *p++ = '\n';
        out(w, o, (size_t)(p - o));
    }
    /* records */
    for (r = 0; r < recs->n; r++) {
        const swm_record_t *x = &recs->rec[r];
        const uint64_t i = x->read;
        const uint32_t nl = name_bytes(names, name_lens, i), L = lens[i];
        const uint32_t c = contig_of(w, x->pos);
        char *o = need(w, (size_t)nl + w->name_len[c] + 2 * (size_t)L + x->cigar_len + 128);
        if (!o) return SWM_ENOMEM;
        char *p = put(o, names[i], nl);
        *p++ = '\t';
        if (x->strand == 0) p = x->primary ? PUTS(p, "0") : PUTS(p, "256");
        else p = x->primary ? PUTS(p, "16") : PUTS(p, "272");
        *p++ = '\t';
        p = put(p, w->name[c], w->name_len[c]);
        *p++ = '\t';
        p = put_int(p, (int)(uint32_t)(x->pos - w->start[c]));
        p = PUTS(p, "\t255\t");
        p = put(p, recs->cigar + x->cigar_off, x->cigar_len);
        p = PUTS(p, "\t*\t0\t0\t");
        if (x->primary) {
            const char *sq = seqs + i * stride, *ql = quals ? quals + i * stride : NULL;
            if (x->strand) {
                if (L > w->rcap) {
                    char *a = (char *)realloc(w->rseq, L), *b = a ? (char *)realloc(w->rqual, L) : NULL;
                    if (a) w->rseq = a;
                    if (b) w->rqual = b;
                    if (!a || !b) return SWM_ENOMEM;
                    w->rcap = L;
                }
                for (uint32_t k = 0; k < L; k++) {           /* reverse_read (mapper_mpe.c:342-366) */
                    const char b = sq[k];
                    char v = b;
                    if (b == 'A') v = 'T';
                    else if (b == 'C') v = 'G';
                    else if (b == 'G') v = 'C';
                    else if (b == 'T') v = 'A';
                    w->rseq[L - 1 - k] = v;
                    if (ql) w->rqual[L - 1 - k] = ql[k];
                }
                sq = w->rseq;
                if (ql) ql = w->rqual;
            }
            p = put(p, sq, L);
            *p++ = '\t';
            if (ql) p = put(p, ql, L);
            else *p++ = '*';
            *p++ = '\t';
        } else {
            p = PUTS(p, "*\t*\t");
        }
        const char d = (char)('0' + x->dist);
        p = PUTS(p, "NM:i:");
        *p++ = d;
        p = PUTS(p, "\tXM:i:");
        *p++ = d;
        *p++ = '\n';
        out(w, o, (size_t)(p - o));
    }
    return w->err;
}

int swm_sam_close(swm_sam_t *w)
{
    if (!w) return SWM_EINVAL;
    int rc = w->err;
    if (w->gz) {
        if (gzclose(w->gz) != Z_OK && rc == SWM_OK) rc = SWM_EIO;
    } else if (w->f) {
        if (fclose(w->f) != 0 && rc == SWM_OK) rc = SWM_EIO;
    }
    free_writer(w);
    return rc;
}
