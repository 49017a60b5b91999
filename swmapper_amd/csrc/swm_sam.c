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
 *
 * A batch is cut into items of up to 8192 lines (its unmapped reads first,
 * then its records), which worker threads format -- and with gzip compress
 * into one gzip member each -- in parallel; the items are written in order.
 * Concatenated members are one gzip stream (RFC 1952), so the file reads
 * back as the same text whatever the thread count.  The reference writes
 * through one gzFile at level 1 (mapper_mpe.c:866).
 */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <pthread.h>
#include <unistd.h>
#include <zlib.h>

#include "swm.h"

struct swm_sam {
    FILE *f;
    int level;              /* 0 plain text, 1..9 gzip members */
    int threads;
    int err;
    uint32_t n_contigs;
    uint32_t *start;        /* n_contigs + 1 cumulative starts (u32, as the reference's ref_len) */
    char **name;            /* n_contigs + 1 (the last one empty) */
    uint32_t *name_len;     /* bytes written for RNAME (the .ref_info name-length line) */
};

/* A growable byte buffer. */
typedef struct {
    char *p;
    size_t n, cap;
} buf_t;

static char *buf_room(buf_t *b, size_t more)
{
    if (b->n + more > b->cap) {
        size_t c = b->cap ? b->cap : 65536;
        while (c < b->n + more) c *= 2;
        char *q = (char *)realloc(b->p, c);
        if (!q) return NULL;
        b->p = q;
        b->cap = c;
    }
    return b->p + b->n;
}

/* One gzip member (RFC 1952) of the text: concatenated members are one
 * valid gzip stream, so chunks compress in parallel. */
static int gzip_member(const char *src, size_t n, int level, buf_t *dst)
{
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (deflateInit2(&zs, level, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK) return SWM_ENOMEM;
    const uLong bound = deflateBound(&zs, (uLong)n) + 64;
    char *o = buf_room(dst, bound);
    if (!o) { deflateEnd(&zs); return SWM_ENOMEM; }
    zs.next_in = (Bytef *)src;
    zs.avail_in = (uInt)n;
    zs.next_out = (Bytef *)o;
    zs.avail_out = (uInt)bound;
    const int r = deflate(&zs, Z_FINISH);
    const size_t made = bound - zs.avail_out;
    deflateEnd(&zs);
    if (r != Z_STREAM_END) return SWM_EIO;
    dst->n += made;
    return SWM_OK;
}

static void emit(swm_sam_t *w, const char *p, size_t n)
{
    if (w->err || n == 0) return;
    if (fwrite(p, 1, n, w->f) != n) w->err = SWM_EIO;
}

/* Text out: compressed to one member when gzip is on. */
static void emit_text(swm_sam_t *w, const char *p, size_t n)
{
    if (w->err || n == 0) return;
    if (w->level == 0) { emit(w, p, n); return; }
    buf_t z = {0, 0, 0};
    const int rc = gzip_member(p, n, w->level, &z);
    if (rc) w->err = rc;
    else emit(w, z.p, z.n);
    free(z.p);
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
    w->level = gzip_level;
    long nc = sysconf(_SC_NPROCESSORS_ONLN);
    w->threads = nc < 1 ? 1 : (nc > 16 ? 16 : (int)nc);
    w->f = fopen(path, "wb");
    if (!w->f) { free_writer(w); return SWM_EIO; }
    /* header (mapper_mpe.c:315-341) */
    buf_t h = {0, 0, 0};
    static const char hd[] = "@HD\tVN:1.4\tSO:unsorted\n";
    static const char pg[] = "@PG\tID:XXX\tVN:1.0\tPN:XXX\tCL:\"CMDLINE\"\n";
    char *o = buf_room(&h, sizeof(hd));
    if (o) { memcpy(o, hd, sizeof(hd) - 1); h.n += sizeof(hd) - 1; }
    for (uint32_t i = 0; i < w->n_contigs && o; i++) {
        const size_t nl = strlen(w->name[i]);
        o = buf_room(&h, nl + 64);
        if (!o) break;
        char *p = PUTS(o, "@SQ\tSN:");
        p = put(p, w->name[i], nl);
        p += sprintf(p, "\tLN:%u\n", (unsigned)(w->start[i + 1] - w->start[i]));
        h.n += (size_t)(p - o);
    }
    if (o && (o = buf_room(&h, sizeof(pg)))) { memcpy(o, pg, sizeof(pg) - 1); h.n += sizeof(pg) - 1; }
    if (!o) { free(h.p); swm_sam_close(w); return SWM_ENOMEM; }
    emit_text(w, h.p, h.n);
    free(h.p);
    if (w->err) { rc = w->err; swm_sam_close(w); return rc; }
    *outp = w;
    return SWM_OK;
}

int swm_sam_set_threads(swm_sam_t *w, int n)
{
    if (!w || n < 1 || n > 256) return SWM_EINVAL;
    w->threads = n;
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

/* ---------------------------------------------------------------- batches */

typedef struct {
    const swm_sam_t *w;
    uint64_t n;
    const char *const *names;
    const uint32_t *name_lens;
    const char *seqs, *quals;
    uint64_t stride;
    const uint32_t *lens;
    const swm_records_t *recs;
    const uint8_t *mapped;      /* per read: has a record */
} batch_t;

#define ITEM_LINES 8192

/* A work item: unmapped lines of reads [a, b) (kind 0) or the lines of
 * records [a, b) (kind 1), formatted and (with gzip) compressed alone. */
typedef struct {
    int kind;
    uint64_t a, b;
    buf_t text, gz;
    int rc;
} item_t;

static uint32_t name_bytes(const batch_t *B, uint64_t i)
{
    return B->name_lens ? B->name_lens[i] : (uint32_t)strlen(B->names[i]);
}

static int fmt_unmapped(const batch_t *B, uint64_t a, uint64_t b, buf_t *t)
{
    for (uint64_t i = a; i < b; i++) {
        if (B->mapped[i]) continue;
        const uint32_t nl = name_bytes(B, i), L = B->lens[i];
        char *o = buf_room(t, (size_t)nl + 2 * (size_t)L + 64);
        if (!o) return SWM_ENOMEM;
        char *p = put(o, B->names[i], nl);
        p = PUTS(p, "\t4\t*\t0\t0\t*\t*\t0\t0\t");
        p = put(p, B->seqs + i * B->stride, L);
        *p++ = '\t';
        if (B->quals) p = put(p, B->quals + i * B->stride, L);
        else *p++ = '*';
        *p++ = '\n';
        t->n += (size_t)(p - o);
    }
    return SWM_OK;
}

static int fmt_records(const batch_t *B, uint64_t a, uint64_t b, buf_t *t)
{
    const swm_sam_t *w = B->w;
    const swm_records_t *recs = B->recs;
    char *rseq = NULL, *rqual = NULL;
    size_t rcap = 0;
    int rc = SWM_OK;
    for (uint64_t r = a; r < b && rc == SWM_OK; r++) {
        const swm_record_t *x = &recs->rec[r];
        const uint64_t i = x->read;
        const uint32_t nl = name_bytes(B, i), L = B->lens[i];
        const uint32_t c = contig_of(w, x->pos);
        char *o = buf_room(t, (size_t)nl + w->name_len[c] + 2 * (size_t)L + x->cigar_len + 128);
        if (!o) { rc = SWM_ENOMEM; break; }
        char *p = put(o, B->names[i], nl);
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
            const char *sq = B->seqs + i * B->stride, *ql = B->quals ? B->quals + i * B->stride : NULL;
            if (x->strand) {
                if (L > rcap) {
                    char *na = (char *)realloc(rseq, L), *nb = na ? (char *)realloc(rqual, L) : NULL;
                    if (na) rseq = na;
                    if (nb) rqual = nb;
                    if (!na || !nb) { rc = SWM_ENOMEM; break; }
                    rcap = L;
                }
                for (uint32_t k = 0; k < L; k++) {           /* reverse_read (mapper_mpe.c:342-366) */
                    const char bc = sq[k];
                    char v = bc;
                    if (bc == 'A') v = 'T';
                    else if (bc == 'C') v = 'G';
                    else if (bc == 'G') v = 'C';
                    else if (bc == 'T') v = 'A';
                    rseq[L - 1 - k] = v;
                    if (ql) rqual[L - 1 - k] = ql[k];
                }
                sq = rseq;
                if (ql) ql = rqual;
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
        t->n += (size_t)(p - o);
    }
    free(rseq);
    free(rqual);
    return rc;
}

static void run_item(const batch_t *B, item_t *it)
{
    it->rc = it->kind == 0 ? fmt_unmapped(B, it->a, it->b, &it->text) : fmt_records(B, it->a, it->b, &it->text);
    if (it->rc == SWM_OK && B->w->level > 0 && it->text.n) {
        it->rc = gzip_member(it->text.p, it->text.n, B->w->level, &it->gz);
        free(it->text.p);
        it->text.p = NULL;
        it->text.n = it->text.cap = 0;
    }
}

typedef struct {
    const batch_t *B;
    item_t *items;
    uint64_t n_items;
    uint64_t next;
    pthread_mutex_t mu;
} pool_t;

static void *pool_worker(void *arg)
{
    pool_t *P = (pool_t *)arg;
    for (;;) {
        pthread_mutex_lock(&P->mu);
        const uint64_t k = P->next < P->n_items ? P->next++ : P->n_items;
        pthread_mutex_unlock(&P->mu);
        if (k >= P->n_items) break;
        run_item(P->B, &P->items[k]);
    }
    return NULL;
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
    uint8_t *mapped = (uint8_t *)calloc(n ? n : 1, 1);
    if (!mapped) return SWM_ENOMEM;
    for (uint64_t r = 0; r < recs->n; r++) mapped[recs->rec[r].read] = 1;   /* put_cigar's unmap: cigar_size == 0 */
    batch_t B = {w, n, names, name_lens, seqs, quals, stride, lens, recs, mapped};
    /* unmapped reads first (batch order), then the records */
    const uint64_t nu = (n + ITEM_LINES - 1) / ITEM_LINES, nr = (recs->n + ITEM_LINES - 1) / ITEM_LINES;
    const uint64_t n_items = nu + nr;
    item_t *items = (item_t *)calloc(n_items ? n_items : 1, sizeof(item_t));
    if (!items) { free(mapped); return SWM_ENOMEM; }
    for (uint64_t k = 0; k < nu; k++) {
        items[k].kind = 0;
        items[k].a = k * ITEM_LINES;
        items[k].b = (k + 1) * ITEM_LINES < n ? (k + 1) * ITEM_LINES : n;
    }
    for (uint64_t k = 0; k < nr; k++) {
        items[nu + k].kind = 1;
        items[nu + k].a = k * ITEM_LINES;
        items[nu + k].b = (k + 1) * ITEM_LINES < recs->n ? (k + 1) * ITEM_LINES : recs->n;
    }
    /* items in groups, so that at most a group's output is held at once;
     * each group is formatted (and compressed) in parallel, then written in order */
    const uint64_t group = (uint64_t)w->threads * 4;
    int rc = SWM_OK;
    for (uint64_t g0 = 0; g0 < n_items && rc == SWM_OK; g0 += group) {
        const uint64_t g1 = g0 + group < n_items ? g0 + group : n_items;
        pool_t P;
        P.B = &B;
        P.items = items + g0;
        P.n_items = g1 - g0;
        P.next = 0;
        pthread_mutex_init(&P.mu, NULL);
        int nt = w->threads;
        if ((uint64_t)nt > P.n_items) nt = (int)P.n_items;
        pthread_t th[256];
        int started = 0;
        for (int t = 1; t < nt; t++)
            if (pthread_create(&th[started], NULL, pool_worker, &P) == 0) started++;
        pool_worker(&P);
        for (int t = 0; t < started; t++) pthread_join(th[t], NULL);
        pthread_mutex_destroy(&P.mu);
        for (uint64_t k = g0; k < g1; k++) {
            item_t *it = &items[k];
            if (it->rc && rc == SWM_OK) rc = it->rc;
            if (rc == SWM_OK) {
                if (w->level > 0) emit(w, it->gz.p, it->gz.n);
                else emit(w, it->text.p, it->text.n);
            }
            free(it->text.p);
            free(it->gz.p);
        }
    }
    free(items);
    free(mapped);
    if (rc != SWM_OK && w->err == SWM_OK) w->err = rc;
    return w->err;
}

int swm_sam_close(swm_sam_t *w)
{
    if (!w) return SWM_EINVAL;
    int rc = w->err;
    if (w->f && fclose(w->f) != 0 && rc == SWM_OK) rc = SWM_EIO;
    free_writer(w);
    return rc;
}

/* ---- the reference's output buffers ---- */

/* put_cigar (mapper_cpe.c:1437-1507) fills, per batch and block b:
 *   ref_pos_buffer[b * (MAX_CIGAR + 2)]      cur_read = 2 * read + strand
 *                  [.. + 1]                  cigar_num | main_flag << 31
 *                  [.. + 2 .. + 2 + num)     the records' locations
 *   cigar_buffer[b * MAX_CIGAR * CIGAR_LEN]  per record: distance byte,
 *                                            text-length byte, text
 * and, for a read with no record on either strand (cigar_size == 0 at the
 * reverse strand's last flush, :1447-1460), cur_read of that reverse strand
 * into cigar_num[].  sam_format_2 (mapper_mpe.c:426-687) reads them with
 * the same strides (LOCATIONS = 258, CIGARS_PER_BLOCK * CIGAR_AVG_LEN =
 * 4096).  Blocks come in record order, unmapped reads in batch order. */
int swm_records_to_ref_blocks(const swm_records_t *recs, uint64_t n_reads, swm_ref_blocks_t *out)
{
    if (!recs || !out || (recs->n && (!recs->rec || !recs->cigar || !recs->block || !recs->n_blocks)))
        return SWM_EINVAL;
    memset(out, 0, sizeof(*out));
    const uint64_t nb = recs->n ? recs->n_blocks : 0;
    const uint64_t pos_stride = SWM_REF_MAX_CIGAR + 2, text_stride = (uint64_t)SWM_REF_MAX_CIGAR * SWM_REF_CIGAR_LEN;
    /* the blocks must tile the records: ascending starts from record 0, each
     * within one (read, strand) and within put_cigar's limits */
    for (uint64_t b = 0; b < nb; b++) {
        const uint64_t r0 = recs->block[b], r1 = b + 1 < nb ? recs->block[b + 1] : recs->n;
        if ((b == 0 && r0 != 0) || r1 <= r0 || r1 > recs->n || r1 - r0 > SWM_REF_MAX_CIGAR) return SWM_EINVAL;
        uint64_t bytes = 0;
        for (uint64_t r = r0; r < r1; r++) {
            const swm_record_t *x = &recs->rec[r];
            if (x->read >= n_reads || x->read != recs->rec[r0].read || x->strand != recs->rec[r0].strand ||
                (uint64_t)x->cigar_off + x->cigar_len > recs->text_bytes || (x->primary && r != r0))
                return SWM_EINVAL;
            bytes += 2u + x->cigar_len;
        }
        /* a block past 4096 bytes overruns put_cigar's buffer in the reference */
        if (bytes > text_stride) return SWM_EINVAL;
    }
    uint8_t *mapped = (uint8_t *)calloc(n_reads ? n_reads : 1, 1);
    if (!mapped) return SWM_ENOMEM;
    for (uint64_t r = 0; r < recs->n; r++) mapped[recs->rec[r].read] = 1;
    uint64_t nu = 0;
    for (uint64_t i = 0; i < n_reads; i++) nu += !mapped[i];
    out->ref_pos = (uint32_t *)calloc(nb ? nb * pos_stride : 1, sizeof(uint32_t));
    out->cigar = (char *)calloc(nb ? nb * text_stride : 1, 1);
    out->unmapped = (uint32_t *)malloc((nu ? nu : 1) * sizeof(uint32_t));
    if (!out->ref_pos || !out->cigar || !out->unmapped) {
        free(mapped);
        swm_ref_blocks_free(out);
        return SWM_ENOMEM;
    }
    for (uint64_t i = 0, u = 0; i < n_reads; i++)
        if (!mapped[i]) out->unmapped[u++] = (uint32_t)(2 * i + 1);
    for (uint64_t b = 0; b < nb; b++) {
        const uint64_t r0 = recs->block[b], r1 = b + 1 < nb ? recs->block[b + 1] : recs->n;
        const swm_record_t *h = &recs->rec[r0];
        uint32_t *p = out->ref_pos + b * pos_stride;
        char *t = out->cigar + b * text_stride;
        p[0] = 2u * h->read + h->strand;
        p[1] = (uint32_t)(r1 - r0) | (h->primary ? 1u << 31 : 0u);
        for (uint64_t r = r0; r < r1; r++) {
            const swm_record_t *x = &recs->rec[r];
            p[2 + (r - r0)] = x->pos;
            *t++ = (char)x->dist;
            *t++ = (char)x->cigar_len;
            memcpy(t, recs->cigar + x->cigar_off, x->cigar_len);
            t += x->cigar_len;
        }
    }
    free(mapped);
    out->n_blocks = nb;
    out->n_unmapped = nu;
    return SWM_OK;
}

void swm_ref_blocks_free(swm_ref_blocks_t *b)
{
    if (!b) return;
    free(b->ref_pos);
    free(b->cigar);
    free(b->unmapped);
    memset(b, 0, sizeof(*b));
}
