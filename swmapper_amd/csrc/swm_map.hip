/*
 * swm_map.hip -- MI355X mapping hot path: seed lookup, candidate batches,
 * 15-mer filter, banded Myers verification and record emission.
 *
 * The path replaces find_positions (mapper_cpe.c:2592-2663) behind
 * swm_map_batch*.  One launch per stage on the context's stream:
 *
 *   k_seed     one thread per (read, strand): read encoding, 12-mer hash
 *              offset gathers, long-seed ranking, residuals
 *              (mapper_cpe.c:1246-1401, 2451-2499)
 *   k_plan     one thread per (read, strand): the batch/heap state machine of
 *              alignment_rmdup (mapper_cpe.c:2124-2426) with the 15-mer filter
 *              (:2090-2123) applied as positions stream in, then the per-batch
 *              sorted unique list (block_heap_sort :1999-2089)
 *   k_compact  one thread per (read, strand): dense candidate list
 *   k_myers    one candidate per lane: window fetch + banded Myers
 *              (semantics mapper_cpe.c:519-660 == :668-976)
 *   k_emit     one thread per read: Hamming check / traceback, duplicate
 *              rule and flush state machine, per-read reduction
 *              (mapper_cpe.c:229-517, 1437-1507, 1790, 1858)
 */
#include "swm_internal.h"

#include <hipcub/hipcub.hpp>
#include <stdlib.h>
#include <string.h>

namespace {

/* ---------------------------------------------------------------- helpers */

__device__ __forceinline__ uint32_t shl32(uint32_t x, int s) { return (s < 0 || s >= 32) ? 0u : (x << s); }
__device__ __forceinline__ uint32_t shr32(uint32_t x, int s) { return (s < 0 || s >= 32) ? 0u : (x >> s); }

/* Plane word (0 = low bit plane, 1 = high) of the genome starting at absolute
 * base g.  The device copy of .base_2bits carries SWM_WIN_PAIRS_PAD zero
 * pairs after the file, so windows past the end read 0; bases before 0 read 0. */
__device__ __forceinline__ uint32_t genome_word(const uint32_t *__restrict__ b2, int64_t g, int plane)
{
    if (g <= -32) return 0u;
    if (g < 0) return shl32(b2[plane], (int)(-g));
    uint64_t wi = (uint64_t)g >> 5;
    uint32_t sh = (uint32_t)g & 31u;
    uint32_t a = b2[2 * wi + plane], b = b2[2 * wi + 2 + plane];
    return __builtin_amdgcn_alignbit(b, a, sh);
}

__device__ __forceinline__ char strand_char(const char *__restrict__ rd, uint32_t L, int strand, uint32_t t)
{
    if (!strand) return rd[t];
    char c = rd[L - 1 - t];
    switch (c) {                       /* reverse_read, mapper_cpe.c:2429-2450 */
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return c;
    }
}

/* pos_start / adjust_ref_len (mapper_cpe.c:1403-1435). */
__device__ __forceinline__ uint32_t pos_start(uint64_t ref_length, uint64_t e, uint32_t base_pos, uint32_t L,
                                              uint32_t *ref_len_out)
{
    uint64_t shift = e;
    uint64_t ref_len = (uint64_t)L + 2 * e;
    if (base_pos <= shift) {
        shift = base_pos;
        ref_len -= (e - base_pos);
    }
    int64_t front = (int64_t)ref_length - (int64_t)((uint64_t)base_pos - e);
    int64_t back = (int64_t)((uint64_t)base_pos + L + e) - (int64_t)ref_length;
    if (front > back) {
        if (back >= 0) ref_len -= (uint64_t)(back + 1);
    } else {
        shift = e - (uint64_t)front - 1;
        ref_len -= (uint64_t)front;
    }
    uint32_t rl = (uint32_t)ref_len;
    *ref_len_out = rl;
    if ((uint64_t)rl < (uint64_t)L - e) return 0xffffffffu;
    return (uint32_t)((uint64_t)base_pos - shift);
}

/* ------------------------------------------------------------------ k_seed */

struct SeedOut {
    uint32_t *words;    /* [nrs][32]: 16 interleaved 2-bit words, 8 plane pairs */
    swm_sel_t *sel;     /* [nrs][seln] */
    uint32_t *nsel;     /* [nrs] selected 12-mers (4 per selected long seed) */
    uint64_t *npos;     /* [nrs + 1] positions over the selected seeds */
    uint32_t *nchunk;   /* [nrs + 1] sum of ceil(cnt / 32) */
};

/* One base of get_seed_from_reads' encoder (mapper_cpe.c:1258-1305).  cp is
 * the interleaved/plane code. */
__device__ __forceinline__ void enc_base(uint32_t (&w)[32], uint32_t &enc, uint64_t rp, uint32_t cp)
{
    enc |= cp << ((rp % 16) * 2);
    w[16 + (rp >> 5) * 2] |= (cp & 1u) << (rp & 0x1f);
    w[16 + (rp >> 5) * 2 + 1] |= (cp >> 1) << (rp & 0x1f);
    if (rp % 16 == 15) {
        w[rp / 16] = enc;
        enc = 0;
    }
}

__device__ __forceinline__ uint32_t plane_code(char b)
{
    switch (b) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return 1;      /* non-ACGT encodes as C (mapper_cpe.c:1295-1299, Q15) */
    }
}

__global__ void __launch_bounds__(128) k_seed(const char *__restrict__ reads, uint64_t stride,
                                              const uint32_t *__restrict__ lens, uint32_t nrs, uint32_t k,
                                              uint32_t e, uint32_t seln, const uint32_t *__restrict__ hash_offset,
                                              SeedOut o)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrs) return;
    const uint32_t read = r >> 1;
    const int strand = (int)(r & 1u);
    const uint32_t L = lens[read];
    const char *rd = reads + (uint64_t)read * stride;

    uint32_t w[32];
#pragma unroll
    for (int j = 0; j < 32; j++) w[j] = 0;
    uint32_t enc = 0, seed = 0;
    const uint32_t seed_len = k + SWM_SEED_STEP - 1;
    const uint32_t n_long = L / seed_len;
    const uint32_t seeds_size = n_long * SWM_SEED_STEP;
    uint32_t off[64], cnt[64];

    /* choose_seed (mapper_cpe.c:1365-1384) driving get_seed_from_reads
     * (:1246-1364) seed by seed, including its tail branch (Q3). */
    for (uint32_t sn = 0; sn < seeds_size; sn++) {
        uint64_t i = (sn % SWM_SEED_STEP == 0) ? 0 : k - 1;
        uint64_t roff = (uint64_t)(sn / SWM_SEED_STEP) * seed_len;
        uint64_t soff = sn % SWM_SEED_STEP;
        uint64_t rp = roff + soff + i;
        for (; i < k; i++) {
            rp = roff + soff + i;
            char b = strand_char(rd, L, strand, (uint32_t)rp);
            uint32_t cs;
            switch (b) {
            case 'A': cs = 0; break;
            case 'C': cs = 1; break;
            case 'G': cs = 2; break;
            case 'T': cs = 3; break;
            default: cs = (uint32_t)(i % 4); break;
            }
            seed = (seed >> 2) | (cs << ((k - 1) * 2));
            enc_base(w, enc, rp, plane_code(b));
        }
        if ((uint64_t)L - rp < k) {
            for (; rp < L; rp++) enc_base(w, enc, rp, plane_code(strand_char(rd, L, strand, (uint32_t)rp)));
            w[(rp - 1) / 16] = enc;
        }
        uint32_t a = hash_offset[seed], b2 = hash_offset[seed + 1];
        off[sn] = a;
        cnt[sn] = b2 - a;
    }
    /* long-seed totals and the exchange sort (mapper_cpe.c:1385-1399, Q4) */
    uint32_t icount[16], iindex[16];
#pragma unroll
    for (int j = 0; j < 16; j++) { icount[j] = 0; iindex[j] = (uint32_t)j; }
    for (uint32_t j = 0; j < seeds_size; j++) icount[j / SWM_SEED_STEP] += cnt[j];
    for (uint32_t j = 0; j < n_long; j++)
        for (uint32_t q = j + 1; q < n_long; q++)
            if (icount[j] > icount[q]) {
                uint32_t tc = icount[j]; icount[j] = icount[q]; icount[q] = tc;
                uint32_t ti = iindex[j]; iindex[j] = iindex[q]; iindex[q] = ti;
            }

    /* selected 12-mers: the e+1 rarest long seeds, fewer if the read has
     * fewer (Q5); residuals as set_read_long_seeds (mapper_cpe.c:2451-2499) */
    uint32_t n_sel_long = (e + 1 < n_long) ? e + 1 : n_long;
    uint32_t nsel = n_sel_long * SWM_SEED_STEP;
    const uint32_t seed_mask = (1u << seed_len) - 1;
    uint64_t npos = 0;
    uint32_t nchunk = 0;
    swm_sel_t *so = o.sel + (uint64_t)r * seln;
    for (uint32_t i = 0; i < nsel; i++) {
        uint32_t m = iindex[i / SWM_SEED_STEP];
        uint32_t q = i % SWM_SEED_STEP;
        uint32_t si = m * SWM_SEED_STEP + q;
        uint32_t so_ = m * seed_len;
        uint32_t j = so_ >> 5, sh = so_ & 31u;
        uint32_t lo_next = (j + 1 < 8) ? w[16 + (j + 1) * 2] : 0u;
        uint32_t hi_next = (j + 1 < 8) ? w[16 + (j + 1) * 2 + 1] : 0u;
        uint32_t l_read = (shr32(w[16 + j * 2], (int)sh) | shl32(lo_next, 32 - (int)sh)) & seed_mask;
        uint32_t h_read = (shr32(w[16 + j * 2 + 1], (int)sh) | shl32(hi_next, 32 - (int)sh)) & seed_mask;
        uint32_t fmask = (1u << q) - 1;
        uint32_t rl = ((l_read & fmask) | ((l_read >> (k + q)) << q)) & 7u;
        uint32_t rh = ((h_read & fmask) | ((h_read >> (k + q)) << q)) & 7u;
        swm_sel_t s;
        s.off = off[si];
        s.cnt = cnt[si];
        s.meta = (m * seed_len + q) | (rl << 16) | (rh << 19) | (q << 22);
        so[i] = s;
        npos += s.cnt;
        nchunk += (s.cnt + SWM_CHUNK - 1) / SWM_CHUNK;
    }
    uint32_t *wo = o.words + (uint64_t)r * 32;
#pragma unroll
    for (int j = 0; j < 32; j++) wo[j] = w[j];
    o.nsel[r] = nsel;
    o.npos[r] = npos;
    o.nchunk[r] = nchunk + 1;   /* batches <= chunks + 1 (an empty first batch) */
}

/* ------------------------------------------------------------------ k_plan */

struct PlanIO {
    const swm_sel_t *sel;
    const uint32_t *nsel;
    const uint64_t *surv_off;   /* exclusive scan of npos */
    const uint32_t *bd_off;     /* exclusive scan of nchunk */
    uint32_t *surv;             /* survivor / candidate values */
    uint2 *bdesc;               /* per batch: (start relative to surv_off, unique count) */
    uint32_t *nbatch;           /* [nrs] */
    uint32_t *nuniq;            /* [nrs + 1] */
};

__device__ __forceinline__ void sort_unique(uint32_t *a, uint32_t n, uint32_t *out_n)
{
    if (n <= 1) { *out_n = n; return; }
    if (n <= 48) {
        for (uint32_t i = 1; i < n; i++) {
            uint32_t v = a[i];
            uint32_t j = i;
            while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; j--; }
            a[j] = v;
        }
    } else {                           /* heapsort in place */
        for (uint32_t start = n / 2; start-- > 0;) {
            uint32_t root = start;
            for (;;) {
                uint32_t c = 2 * root + 1;
                if (c >= n) break;
                if (c + 1 < n && a[c + 1] > a[c]) c++;
                if (a[root] < a[c]) { uint32_t t = a[root]; a[root] = a[c]; a[c] = t; root = c; }
                else break;
            }
        }
        for (uint32_t end = n - 1; end > 0; end--) {
            uint32_t t = a[0]; a[0] = a[end]; a[end] = t;
            uint32_t root = 0;
            for (;;) {
                uint32_t c = 2 * root + 1;
                if (c >= end) break;
                if (c + 1 < end && a[c + 1] > a[c]) c++;
                if (a[root] < a[c]) { uint32_t t2 = a[root]; a[root] = a[c]; a[c] = t2; root = c; }
                else break;
            }
        }
    }
    uint32_t u = 1;
    for (uint32_t i = 1; i < n; i++)
        if (a[i] != a[u - 1]) a[u++] = a[i];
    *out_n = u;
}

__global__ void __launch_bounds__(64) k_plan(uint32_t nrs, const uint32_t *__restrict__ positions,
                                             const uint16_t *__restrict__ lseeds, PlanIO io)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrs) return;
    const uint32_t nsel = io.nsel[r];
    const swm_sel_t *sel = io.sel + (uint64_t)r * 64;
    uint32_t *surv = io.surv + io.surv_off[r];
    uint2 *bd = io.bdesc + io.bd_off[r];

    uint32_t pcount[64];
    uint32_t hloc[66], hind[66];      /* seed heap, 1-based (mapper_cpe.c:2149) */
    uint32_t hsize = 0;
    uint32_t nsurv = 0, bstart = 0, batch = 0, pos_size = 0, blocks = 0;
    bool first_pending = true;        /* Q7: the batch's first position is never kept */

    /* positions [g, g+len) of selected seed i enter the batch: 15-mer filter
     * (long_seeds_filter, :2106-2117) and read-start adjustment
     * (adjust_ref_pos_rmdup, :1903-1905) */
    auto take = [&](uint32_t g, uint32_t len, uint32_t meta) {
        const uint32_t bo = meta & 0xffffu, rl = (meta >> 16) & 7u, rh = (meta >> 19) & 7u, q = (meta >> 22) & 3u;
        for (uint32_t j = 0; j < len; j++) {
            uint32_t ls = lseeds[g + j];
            if (first_pending) { first_pending = false; continue; }
            uint32_t res = (((ls >> (3 - q)) ^ rl) | ((ls >> (9 - q)) ^ rh)) & 7u;
            if (res == 0) {
                uint32_t p = positions[g + j];
                surv[nsurv++] = p > bo ? p - bo : 0u;
            }
        }
    };
    auto heap_insert = [&](uint32_t loc, uint32_t ind) {      /* :1919-1935 */
        hsize++;
        hloc[hsize] = loc; hind[hsize] = ind;
        uint32_t i = hsize;
        while (i > 1) {
            uint32_t ti = i / 2;
            if (hloc[ti] > hloc[i]) {
                uint32_t t = hloc[ti]; hloc[ti] = hloc[i]; hloc[i] = t;
                t = hind[ti]; hind[ti] = hind[i]; hind[i] = t;
            } else break;
            i = ti;
        }
    };
    auto heap_pop_insert = [&](uint32_t loc, bool del) {      /* :1936-1962 (Q6) */
        if (del) { hloc[1] = hloc[hsize]; hind[1] = hind[hsize]; hsize--; }
        else hloc[1] = loc;
        uint32_t i = 1;
        while (i * 2 < hsize) {
            uint32_t ti = i * 2;
            if (i * 2 + 1 <= hsize && hloc[i * 2] > hloc[i * 2 + 1]) ti = i * 2 + 1;
            if (hloc[ti] < hloc[i]) {
                uint32_t t = hloc[ti]; hloc[ti] = hloc[i]; hloc[i] = t;
                t = hind[ti]; hind[ti] = hind[i]; hind[i] = t;
            } else break;
            i = ti;
        }
    };

    /* phase 1 (:2194-2236) */
    for (uint32_t i = 0; i < nsel; i++) {
        swm_sel_t s = sel[i];
        uint32_t get = s.cnt > SWM_CHUNK ? SWM_CHUNK : s.cnt;
        if (get > 0) { take(s.off, get, s.meta); blocks++; }
        pos_size += get;
        pcount[i] = get;
    }
    for (uint32_t i = 0; i < nsel; i++) {                     /* :2244-2252 */
        swm_sel_t s = sel[i];
        if (pcount[i] != s.cnt) heap_insert(positions[s.off + pcount[i] - 1], i);
    }
    for (;;) {                                                /* :2261-2419 */
        if (hsize > 0 && pos_size + SWM_CHUNK <= SWM_BATCH_POS && blocks < SWM_MAX_BLOCKS) {
            uint32_t i = hind[1];
            swm_sel_t s = sel[i];
            uint32_t rem = s.cnt - pcount[i];
            uint32_t get = rem > SWM_CHUNK ? SWM_CHUNK : rem;
            take(s.off + pcount[i], get, s.meta);
            pos_size += get;
            pcount[i] += get;
            blocks++;
            heap_pop_insert(positions[s.off + pcount[i] - 1], pcount[i] == s.cnt);
        }
        if (pos_size + SWM_CHUNK > SWM_BATCH_POS || hsize == 0 || blocks >= SWM_MAX_BLOCKS) {
            uint32_t u = 0;
            sort_unique(surv + bstart, nsurv - bstart, &u);
            bd[batch] = make_uint2(bstart, u);
            nsurv = bstart + u;
            bstart = nsurv;
            batch++;
            pos_size = 0;
            blocks = 0;
            first_pending = true;
            if (hsize == 0) break;
        }
    }
    io.nbatch[r] = batch;
    io.nuniq[r] = nsurv;
}

/* --------------------------------------------------------------- k_compact */

struct CandSet {
    uint32_t *val;      /* read-start candidate */
    uint32_t *rs;       /* (read << 1) | strand */
    uint32_t *bi;       /* batch << 12 | index within batch */
    int32_t *err;       /* -2: not verified (pos_start rejected) */
    int32_t *loc;
};

__global__ void __launch_bounds__(128) k_compact(uint32_t nrs, const uint64_t *__restrict__ surv_off,
                                                 const uint32_t *__restrict__ surv, const uint32_t *__restrict__ bd_off,
                                                 const uint2 *__restrict__ bdesc, const uint32_t *__restrict__ nbatch,
                                                 const uint32_t *__restrict__ cand_off, CandSet c)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrs) return;
    const uint32_t *s = surv + surv_off[r];
    const uint2 *bd = bdesc + bd_off[r];
    uint32_t o = cand_off[r];
    uint32_t nb = nbatch[r];
    for (uint32_t b = 0; b < nb; b++) {
        uint2 d = bd[b];
        for (uint32_t j = 0; j < d.y; j++) {
            c.val[o] = s[d.x + j];
            c.rs[o] = r;
            c.bi[o] = (b << 12) | j;
            o++;
        }
    }
}

/* ----------------------------------------------------------------- k_myers */

/* Banded Myers for one candidate (the recurrence of mapper_cpe.c:570-631 with
 * the end scan of :634-650).  The window planes are aligned at the window
 * start; at read base t the match mask over band bits j = window base t + j
 * is built from the plane bits (the compiled kernel's form, :845-857). */
#define SWM_NW 9   /* window words: ((254 - 1) >> 5) + 2 */

__device__ __forceinline__ void myers_lane(const uint32_t *__restrict__ rw, uint32_t L, uint32_t e,
                                           uint32_t ref_len, uint32_t (&wl)[SWM_NW], uint32_t (&wh)[SWM_NW],
                                           int32_t *err_out, int32_t *loc_out)
{
    const uint32_t B = 2 * e + 1;
    const uint32_t mask = (1u << B) - 1u;
    uint32_t vp = 0, vn = 0, ones = 0;
    const uint32_t nfull = L >> 5, rem = L & 31u;

#define SWM_STEP(S, R)                                                                     \
    {                                                                                      \
        const int sb = 2 * ((S) & 15);                                                     \
        uint32_t RL = (uint32_t)(((int32_t)((R) << (31 - sb))) >> 31);                     \
        uint32_t RH = (uint32_t)(((int32_t)((R) << (30 - sb))) >> 31);                     \
        uint32_t bl = __builtin_amdgcn_alignbit(l1, l0, (S));                              \
        uint32_t bh = __builtin_amdgcn_alignbit(h1, h0, (S));                              \
        uint32_t eq = ~((bl ^ RL) | (bh ^ RH)) & mask;                                     \
        uint32_t x = eq | vn;                                                              \
        uint32_t d0 = (((x & vp) + vp) ^ vp) | x;                                          \
        uint32_t hn = vp & d0;                                                             \
        uint32_t hp = vn | ~(vp | d0);                                                     \
        x = d0 >> 1;                                                                       \
        vn = x & hp;                                                                       \
        vp = hn | ~(x | hp);                                                               \
        ones += d0 & 1u;                                                                   \
    }

    for (uint32_t c = 0; c < nfull; c++) {
        const uint32_t r0 = rw[2 * c], r1 = rw[2 * c + 1];
        const uint32_t l0 = wl[0], l1 = wl[1], h0 = wh[0], h1 = wh[1];
#pragma unroll
        for (int s = 0; s < 16; s++) SWM_STEP(s, r0)
#pragma unroll
        for (int s = 16; s < 32; s++) SWM_STEP(s, r1)
#pragma unroll
        for (int j = 0; j < SWM_NW - 1; j++) { wl[j] = wl[j + 1]; wh[j] = wh[j + 1]; }
    }
    if (rem) {
        const uint32_t r0 = rw[2 * nfull], r1 = (rem > 16) ? rw[2 * nfull + 1] : 0u;
        const uint32_t l0 = wl[0], l1 = wl[1], h0 = wh[0], h1 = wh[1];
#pragma unroll
        for (int s = 0; s < 16; s++) {
            if ((uint32_t)s < rem) SWM_STEP(s, r0)
        }
#pragma unroll
        for (int s = 16; s < 32; s++) {
            if ((uint32_t)s < rem) SWM_STEP(s, r1)
        }
    }
#undef SWM_STEP
    int err = (int)L - (int)ones;
    const int last_high = (int)ref_len - (int)L;
    const int site = (int)L - 1;
    int location = -1, error = SWM_MAXERROR;
    if (err <= (int)e && err < error) { error = err; location = site; }
    for (int i = 0; i < last_high; i++) {
        err += (int)((vp >> i) & 1u);
        err -= (int)((vn >> i) & 1u);
        if (err <= (int)e && err < error) { error = err; location = site + i + 1; }
    }
    *err_out = error;
    *loc_out = location;
}

__device__ __forceinline__ void load_window(const uint32_t *__restrict__ b2, uint32_t win_start, uint32_t L,
                                            uint32_t (&wl)[SWM_NW], uint32_t (&wh)[SWM_NW])
{
    const uint2 *P = reinterpret_cast<const uint2 *>(b2) + (win_start >> 5);
    const uint32_t sh = win_start & 31u;
    const uint32_t nneed = ((L - 1) >> 5) + 2;
    uint2 pr[SWM_NW + 1];
#pragma unroll
    for (int j = 0; j < SWM_NW + 1; j++) pr[j] = ((uint32_t)j <= nneed) ? P[j] : make_uint2(0u, 0u);
#pragma unroll
    for (int j = 0; j < SWM_NW; j++) {
        wl[j] = __builtin_amdgcn_alignbit(pr[j + 1].x, pr[j].x, sh);
        wh[j] = __builtin_amdgcn_alignbit(pr[j + 1].y, pr[j].y, sh);
    }
}

__global__ void __launch_bounds__(256) k_myers(uint32_t ncand, const uint32_t *__restrict__ lens,
                                               const uint32_t *__restrict__ words, const uint32_t *__restrict__ b2,
                                               uint64_t ref_length, uint32_t e, CandSet c,
                                               unsigned long long *n_verif)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid = false;
    if (i < ncand) {
        const uint32_t rs = c.rs[i];
        const uint32_t L = lens[rs >> 1];
        uint32_t ref_len;
        const uint32_t st = pos_start(ref_length, e, c.val[i], L, &ref_len);
        if (st == 0xffffffffu) {
            c.err[i] = -2;
            c.loc[i] = -1;
        } else {
            uint32_t wl[SWM_NW], wh[SWM_NW];
            load_window(b2, st, L, wl, wh);
            int32_t er, lc;
            myers_lane(words + (uint64_t)rs * 32, L, e, ref_len, wl, wh, &er, &lc);
            c.err[i] = er;
            c.loc[i] = lc;
            valid = true;
        }
    }
    unsigned long long bal = __ballot(valid);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(n_verif, (unsigned long long)__popcll(bal));
}

/* Stage entry: explicit candidates. */
__global__ void __launch_bounds__(256) k_myers_explicit(uint32_t n, const uint32_t *__restrict__ read_words,
                                                        const uint32_t *__restrict__ lens,
                                                        const uint32_t *__restrict__ cread,
                                                        const uint32_t *__restrict__ wstart,
                                                        const uint32_t *__restrict__ rlen,
                                                        const uint32_t *__restrict__ b2, uint32_t e,
                                                        int32_t *__restrict__ err, int32_t *__restrict__ loc)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t rr = cread[i];
    const uint32_t L = lens[rr];
    uint32_t wl[SWM_NW], wh[SWM_NW];
    load_window(b2, wstart[i], L, wl, wh);
    int32_t er, lc;
    myers_lane(read_words + (uint64_t)rr * 16, L, e, rlen[i], wl, wh, &er, &lc);
    err[i] = er;
    loc[i] = lc;
}

/* ------------------------------------------------------------------ k_emit */

__device__ __forceinline__ int ndigits(uint32_t v)
{
    int n = 0;
    while (v) { n++; v /= 10; }
    return n;
}

struct EmitState {
    uint32_t cigar_num, cigar_len, last_loc;
    uint32_t count;
    uint32_t best_pos;
    uint32_t best_dist;
    uint32_t best_strand;
};

__device__ __forceinline__ void emit_record(EmitState &em, uint32_t loc, uint32_t dist, uint32_t strand,
                                            uint32_t text_len)
{
    em.cigar_num++;
    em.cigar_len += text_len;
    em.last_loc = loc;
    if (em.count == 0 || dist < em.best_dist) {
        em.best_pos = loc;
        em.best_dist = dist;
        em.best_strand = strand;
    }
    em.count++;
}

/* align_reads for one passing lane (mapper_cpe.c:243-509). */
__device__ void align_one(EmitState &em, const uint32_t *__restrict__ w, const uint32_t *__restrict__ b2,
                          uint32_t win_start, uint32_t ref_len, uint32_t L, uint32_t e, int error, int location,
                          uint32_t strand)
{
    int start_location = location - (int)L + 1;
    int tmp_err = 0;
    for (uint32_t ii = 0; ii * 32 < L; ii++) {                                   /* Hamming :252-276 */
        uint32_t mask = 0xffffffffu;
        if ((ii + 1) * 32 > L) mask >>= (32 - (L - ii * 32));
        uint32_t l0 = genome_word(b2, (int64_t)win_start + 32 * ii, 0);
        uint32_t l1 = genome_word(b2, (int64_t)win_start + 32 * (ii + 1), 0);
        uint32_t h0 = genome_word(b2, (int64_t)win_start + 32 * ii, 1);
        uint32_t h1 = genome_word(b2, (int64_t)win_start + 32 * (ii + 1), 1);
        uint32_t lref = shr32(l0, start_location) | shl32(l1, 32 - start_location);
        uint32_t href = shr32(h0, start_location) | shl32(h1, 32 - start_location);
        uint32_t x = ((w[16 + 2 * ii + 1] ^ href) | (w[16 + 2 * ii] ^ lref)) & mask;
        tmp_err += __popc(x);
    }
    if (tmp_err == error) {
        emit_record(em, (uint32_t)((uint64_t)win_start + (uint64_t)(int64_t)start_location + 1), (uint32_t)error,
                    strand, 2u + (uint32_t)ndigits(L) + 1u);
        return;
    }
    /* traceback: the recurrence again, keeping D0 and HP (:288-362) */
    uint32_t d0a[SWM_READ_MAX + 2], hpa[SWM_READ_MAX + 2];
    {
        const uint32_t band_length = 2 * e + 1, mask = 1u << (band_length - 1);
        uint32_t pl = genome_word(b2, win_start, 0), ph = genome_word(b2, win_start, 1);
        uint32_t peq[4] = {0, 0, 0, 0};
        for (uint32_t j = 0; j < band_length; j++) peq[((pl >> j) & 1u) | (((ph >> j) & 1u) << 1)] |= 1u << j;
        uint32_t vp = 0, vn = 0, i_bd = 2 * e;
        for (uint32_t t = 0; t < L; t++) {
            uint32_t be = (w[t >> 4] >> ((t & 15) * 2)) & 3u;
            uint32_t x = peq[be] | vn;
            uint32_t d0 = (((x & vp) + vp) ^ vp) | x;
            uint32_t hn = vp & d0;
            uint32_t hp = vn | ~(vp | d0);
            x = d0 >> 1;
            vn = x & hp;
            vp = hn | ~(x | hp);
            d0a[t] = d0;
            hpa[t] = hp;
            peq[0] >>= 1; peq[1] >>= 1; peq[2] >>= 1; peq[3] >>= 1;
            ++i_bd;
            uint32_t gl = genome_word(b2, (int64_t)win_start + (i_bd & ~31u), 0);
            uint32_t gh = genome_word(b2, (int64_t)win_start + (i_bd & ~31u), 1);
            peq[((gl >> (i_bd & 31u)) & 1u) | (((gh >> (i_bd & 31u)) & 1u) << 1)] |= mask;
        }
    }
    const int last_high = (int)ref_len - (int)L;
    int match_site = location;
    const int site = (int)ref_len - last_high - 1;
    int search_site = match_site - site;
    int pre_size = 1;
    char pre_char = 'N';
    int i = (int)L - 1;
    int sum_err = 0;
    const int err = error;
    /* runs in reverse order; only their lengths feed the CIGAR byte count */
    int rsize[SWM_READ_MAX + 8];
    char rchar[SWM_READ_MAX + 8];
    auto refb = [&](int rfi) -> uint32_t {
        int64_t g = (int64_t)win_start + rfi;
        return (genome_word(b2, g, 0) & 1u) | ((genome_word(b2, g, 1) & 1u) << 1);
    };
    auto rdb = [&](int rdi) -> uint32_t { return (w[rdi / 16] >> (2 * (rdi % 16))) & 3u; };
    auto bit = [&](uint32_t v, int s) -> uint32_t { return (s < 0 || s >= 32) ? 0u : ((v >> s) & 1u); };
    int j = 1;
    if (bit(d0a[i], search_site) && refb(match_site) == rdb(i)) {                /* :380-410 */
        i--; pre_size = 1; pre_char = 'M'; match_site--;
    } else if (!bit(d0a[i], search_site) && refb(match_site) != rdb(i)) {
        i--; pre_size = 1; pre_char = 'S'; match_site--; sum_err++;
    } else if (bit(hpa[i], search_site)) {
        i--; search_site++; pre_size = 1; pre_char = 'S'; sum_err++; start_location++;
    } else {
        search_site--; pre_size = 1; pre_char = 'D'; match_site--; sum_err++; start_location--;
    }
    while (i >= 0) {                                                             /* :412-466 */
        if (sum_err == err) break;
        char nc;
        if (bit(d0a[i], search_site) && refb(match_site) == rdb(i)) {
            i--; match_site--; nc = 'M';
        } else if (!bit(d0a[i], search_site) && refb(match_site) != rdb(i)) {
            i--; match_site--; sum_err++; nc = 'S';
        } else if (bit(hpa[i], search_site)) {
            i--; search_site++; sum_err++; start_location++; nc = 'I';
        } else {
            search_site--; match_site--; sum_err++; start_location--; nc = 'D';
        }
        if (pre_char != nc) { rsize[j] = pre_size; rchar[j++] = pre_char; pre_size = 1; pre_char = nc; }
        else pre_size++;
    }
    rsize[j] = pre_size; rchar[j++] = pre_char;                                  /* :468-473 */
    if (i >= 0) { rsize[j] = i + 1; rchar[j++] = 'M'; }
    const uint64_t new_loc = (uint64_t)win_start + (uint64_t)(int64_t)start_location + 1;
    if (em.cigar_num > 0 && (uint64_t)em.last_loc == new_loc) return;          /* :475-477 */
    uint32_t text_len = 2;                                                       /* :479-507 */
    for (j = j - 1; j > 0; j--) {
        if (rchar[j] == 'M' || rchar[j] == 'S') {
            int size_sm = 0;
            while (j > 0 && (rchar[j] == 'M' || rchar[j] == 'S')) { size_sm += rsize[j]; j--; }
            j++;
            text_len += (uint32_t)ndigits((uint32_t)size_sm) + 1u;
        } else {
            text_len += (uint32_t)ndigits((uint32_t)rsize[j]) + 1u;
        }
    }
    emit_record(em, (uint32_t)new_loc, (uint32_t)error, strand, text_len);
}

__global__ void __launch_bounds__(64) k_emit(uint32_t nreads, const uint32_t *__restrict__ lens,
                                             const uint32_t *__restrict__ words, const uint32_t *__restrict__ b2,
                                             uint64_t ref_length, uint32_t e, const uint32_t *__restrict__ cand_off,
                                             CandSet c, swm_read_result_t *__restrict__ out,
                                             unsigned long long *n_records)
{
    uint32_t rd = blockIdx.x * blockDim.x + threadIdx.x;
    if (rd >= nreads) return;
    const uint32_t L = lens[rd];
    EmitState em;
    em.count = 0;
    em.best_pos = 0xffffffffu;
    em.best_dist = 0xffu;
    em.best_strand = 0xffu;
    for (uint32_t strand = 0; strand < 2; strand++) {
        const uint32_t rs = 2 * rd + strand;
        const uint32_t *w = words + (uint64_t)rs * 32;
        em.cigar_num = 0;
        em.cigar_len = 0;
        em.last_loc = 0;
        uint32_t cur_batch = 0xffffffffu, vcount = 0;
        bool open = false;
        auto flush_check = [&]() {                                               /* :1790, :1858 */
            if (em.cigar_num + 8 > 256 || em.cigar_len + 16 * 16 > 256 * 16) {
                em.cigar_num = 0;
                em.cigar_len = 0;
            }
        };
        for (uint32_t ci = cand_off[rs]; ci < cand_off[rs + 1]; ci++) {
            const uint32_t b = c.bi[ci] >> 12;
            if (b != cur_batch) {
                if (open) { flush_check(); open = false; }
                cur_batch = b;
                vcount = 0;
            }
            const int32_t er = c.err[ci];
            if (er == -2) continue;                                              /* not a lane */
            if (open && (vcount & 7u) == 0) { flush_check(); open = false; }
            open = true;
            vcount++;
            if (er <= (int32_t)e) {
                uint32_t ref_len;
                const uint32_t st = pos_start(ref_length, e, c.val[ci], L, &ref_len);
                align_one(em, w, b2, st, ref_len, L, e, er, c.loc[ci], strand);
            }
        }
        if (open) flush_check();
    }
    swm_read_result_t res;
    res.pos = em.count ? em.best_pos : 0xffffffffu;
    res.strand = em.count ? (uint8_t)em.best_strand : (uint8_t)0xff;
    res.dist = em.count ? (uint8_t)em.best_dist : (uint8_t)0xff;
    res.pad = 0;
    res.count = em.count;
    out[rd] = res;
    if (em.count) atomicAdd(n_records, (unsigned long long)em.count);
}

/* --------------------------------------------------------------- host side */

template <typename T>
int dalloc(T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void **)p, n * sizeof(T)) == hipSuccess ? SWM_OK : SWM_ENOMEM;
}

struct Bufs {
    void *ptrs[32];
    int n = 0;
    ~Bufs() { for (int i = 0; i < n; i++) (void)hipFree(ptrs[i]); }
    template <typename T>
    int get(T **p, size_t cnt)
    {
        int rc = dalloc(p, cnt);
        if (rc == SWM_OK) ptrs[n++] = (void *)*p;
        return rc;
    }
};

template <typename T>
int exclusive_scan(T *d_in, T *d_out, size_t n, hipStream_t st)
{
    size_t tmp = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, d_in, d_out, n, st) != hipSuccess) return SWM_EDEVICE;
    void *d_tmp = nullptr;
    if (hipMalloc(&d_tmp, tmp ? tmp : 1) != hipSuccess) return SWM_ENOMEM;
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(d_tmp, tmp, d_in, d_out, n, st);
    (void)hipFree(d_tmp);
    return e == hipSuccess ? SWM_OK : SWM_EDEVICE;
}

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

struct swm_ws {};

/* ---------------------------------------------------------------- C ABI */

extern "C" int swm_ctx_create(uint32_t k, const uint32_t *hash_count, const uint32_t *positions,
                              const uint16_t *pos_lseeds, uint64_t n_pos, const uint32_t *base_2bits,
                              uint64_t n_2bits_words, swm_ctx_t **out)
{
    if (!out || !hash_count || !base_2bits || k < 2 || k > 14 || (n_pos && (!positions || !pos_lseeds)))
        return SWM_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SWM_EDEVICE;
    swm_ctx_t *c = (swm_ctx_t *)calloc(1, sizeof(swm_ctx_t));
    if (!c) return SWM_ENOMEM;
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        free(c);
        return SWM_EDEVICE;
    }
    c->k = k;
    c->n_buckets = 1ull << (2 * k);
    c->n_pos = n_pos;
    c->n_2bits_words = n_2bits_words;
    c->ref_length = 16ull * (n_2bits_words * 4ull);
    /* hash_offset = exclusive scan of hash_count in u32 (mapper_mpe.c:250-255) */
    uint32_t *h_off = (uint32_t *)malloc((c->n_buckets + 1) * sizeof(uint32_t));
    if (!h_off) { swm_ctx_destroy(c); return SWM_ENOMEM; }
    h_off[0] = 0;
    for (uint64_t i = 0; i < c->n_buckets; i++) h_off[i + 1] = h_off[i] + hash_count[i];
    int rc = SWM_OK;
    if (hipMalloc((void **)&c->d_hash_offset, (c->n_buckets + 1) * 4) != hipSuccess ||
        hipMalloc((void **)&c->d_positions, (n_pos + 64) * 4) != hipSuccess ||
        hipMalloc((void **)&c->d_lseeds, (n_pos + 64) * 2) != hipSuccess ||
        hipMalloc((void **)&c->d_base_2bits, (n_2bits_words + 2 * SWM_WIN_PAIRS_PAD) * 4) != hipSuccess)
        rc = SWM_ENOMEM;
    if (rc == SWM_OK) {
        if (hipMemcpy(c->d_hash_offset, h_off, (c->n_buckets + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(c->d_positions, 0, (n_pos + 64) * 4) != hipSuccess ||
            hipMemset(c->d_lseeds, 0, (n_pos + 64) * 2) != hipSuccess ||
            (n_pos && hipMemcpy(c->d_positions, positions, n_pos * 4, hipMemcpyHostToDevice) != hipSuccess) ||
            (n_pos && hipMemcpy(c->d_lseeds, pos_lseeds, n_pos * 2, hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemset(c->d_base_2bits, 0, (n_2bits_words + 2 * SWM_WIN_PAIRS_PAD) * 4) != hipSuccess ||
            hipMemcpy(c->d_base_2bits, base_2bits, n_2bits_words * 4, hipMemcpyHostToDevice) != hipSuccess)
            rc = SWM_EDEVICE;
    }
    free(h_off);
    if (rc != SWM_OK) { swm_ctx_destroy(c); return rc; }
    *out = c;
    return SWM_OK;
}

extern "C" void swm_ctx_destroy(swm_ctx_t *c)
{
    if (!c) return;
    if (c->d_hash_offset) (void)hipFree(c->d_hash_offset);
    if (c->d_positions) (void)hipFree(c->d_positions);
    if (c->d_lseeds) (void)hipFree(c->d_lseeds);
    if (c->d_base_2bits) (void)hipFree(c->d_base_2bits);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    free(c);
}

extern "C" int swm_last_stats(const swm_ctx_t *ctx, swm_stats_t *st)
{
    if (!ctx || !st) return SWM_EINVAL;
    *st = ctx->stats;
    return SWM_OK;
}

static int map_device_impl(swm_ctx_t *ctx, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                           uint64_t n, uint32_t e, swm_read_result_t *d_out)
{
    hipStream_t st = ctx->stream;
    memset(&ctx->stats, 0, sizeof(ctx->stats));
    ctx->stats.n_reads = n;
    if (n == 0) return SWM_OK;
    const uint64_t nrs = 2 * n;
    if (nrs >= 0xffffffffull) return SWM_EINVAL;
    Bufs B;
    SeedOut so;
    int rc;
    hipEvent_t ev[7];
    for (int i = 0; i < 7; i++) SWM_HIP_CHECK(hipEventCreate(&ev[i]));
    struct EvGuard { hipEvent_t *e; ~EvGuard() { for (int i = 0; i < 7; i++) (void)hipEventDestroy(e[i]); } } eg{ev};

    if ((rc = B.get(&so.words, nrs * 32)) || (rc = B.get(&so.sel, nrs * 64)) || (rc = B.get(&so.nsel, nrs)) ||
        (rc = B.get(&so.npos, nrs + 1)) || (rc = B.get(&so.nchunk, nrs + 1)))
        return rc;
    SWM_HIP_CHECK(hipMemsetAsync(so.npos + nrs, 0, sizeof(uint64_t), st));
    SWM_HIP_CHECK(hipMemsetAsync(so.nchunk + nrs, 0, sizeof(uint32_t), st));
    SWM_HIP_CHECK(hipEventRecord(ev[0], st));
    k_seed<<<blocks_for(nrs, 128), 128, 0, st>>>(d_reads, stride, d_lens, (uint32_t)nrs, ctx->k, e, 64,
                                                 ctx->d_hash_offset, so);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(ev[1], st));

    PlanIO pio;
    uint64_t *surv_off;
    uint32_t *bd_off;
    if ((rc = B.get(&surv_off, nrs + 1)) || (rc = B.get(&bd_off, nrs + 1))) return rc;
    if ((rc = exclusive_scan(so.npos, surv_off, nrs + 1, st)) || (rc = exclusive_scan(so.nchunk, bd_off, nrs + 1, st)))
        return rc;
    uint64_t tot_pos = 0;
    uint32_t tot_chunk = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(&tot_pos, surv_off + nrs, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipMemcpyAsync(&tot_chunk, bd_off + nrs, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    pio.sel = so.sel;
    pio.nsel = so.nsel;
    pio.surv_off = surv_off;
    pio.bd_off = bd_off;
    uint32_t *surv;
    uint2 *bdesc;
    uint32_t *nbatch, *nuniq;
    if ((rc = B.get(&surv, tot_pos + 1)) || (rc = B.get(&bdesc, (uint64_t)tot_chunk + 1)) || (rc = B.get(&nbatch, nrs)) ||
        (rc = B.get(&nuniq, nrs + 1)))
        return rc;
    pio.surv = surv;
    pio.bdesc = bdesc;
    pio.nbatch = nbatch;
    pio.nuniq = nuniq;
    SWM_HIP_CHECK(hipMemsetAsync(nuniq + nrs, 0, sizeof(uint32_t), st));
    k_plan<<<blocks_for(nrs, 64), 64, 0, st>>>((uint32_t)nrs, ctx->d_positions, ctx->d_lseeds, pio);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(ev[2], st));

    uint32_t *cand_off;
    if ((rc = B.get(&cand_off, nrs + 1))) return rc;
    if ((rc = exclusive_scan(nuniq, cand_off, nrs + 1, st))) return rc;
    uint32_t ncand = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(&ncand, cand_off + nrs, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    CandSet cs;
    if ((rc = B.get(&cs.val, (uint64_t)ncand + 1)) || (rc = B.get(&cs.rs, (uint64_t)ncand + 1)) ||
        (rc = B.get(&cs.bi, (uint64_t)ncand + 1)) || (rc = B.get(&cs.err, (uint64_t)ncand + 1)) ||
        (rc = B.get(&cs.loc, (uint64_t)ncand + 1)))
        return rc;
    k_compact<<<blocks_for(nrs, 128), 128, 0, st>>>((uint32_t)nrs, surv_off, surv, bd_off, bdesc, nbatch, cand_off,
                                                    cs);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(ev[3], st));

    unsigned long long *counters;
    if ((rc = B.get(&counters, 2))) return rc;
    SWM_HIP_CHECK(hipMemsetAsync(counters, 0, 2 * sizeof(unsigned long long), st));
    if (ncand)
        k_myers<<<blocks_for(ncand, 256), 256, 0, st>>>(ncand, d_lens, so.words, ctx->d_base_2bits,
                                                         ctx->ref_length, e, cs, counters);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(ev[4], st));
    k_emit<<<blocks_for(n, 64), 64, 0, st>>>((uint32_t)n, d_lens, so.words, ctx->d_base_2bits, ctx->ref_length, e,
                                             cand_off, cs, d_out, counters + 1);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(ev[5], st));
    unsigned long long hc[2] = {0, 0};
    SWM_HIP_CHECK(hipMemcpyAsync(hc, counters, sizeof(hc), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    float ms;
    ctx->stats.n_verifications = hc[0];
    ctx->stats.n_records = hc[1];
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]); ctx->stats.ms_seed = ms;
    (void)hipEventElapsedTime(&ms, ev[1], ev[2]); ctx->stats.ms_plan = ms;
    (void)hipEventElapsedTime(&ms, ev[2], ev[3]); ctx->stats.ms_compact = ms;
    (void)hipEventElapsedTime(&ms, ev[3], ev[4]); ctx->stats.ms_myers = ms;
    (void)hipEventElapsedTime(&ms, ev[4], ev[5]); ctx->stats.ms_emit = ms;
    (void)hipEventElapsedTime(&ms, ev[0], ev[5]); ctx->stats.ms_total = ms;
    return SWM_OK;
}

static int check_map_args(const swm_ctx_t *ctx, uint64_t stride, uint32_t e)
{
    if (!ctx) return SWM_EINVAL;
    if (e > SWM_MAX_E) return SWM_EINVAL;              /* band < 32 bits (mapper_cpe.c:686) */
    if (stride < 3 || stride > 256) return SWM_EINVAL; /* -l rule, Q20 */
    return SWM_OK;
}

extern "C" int swm_map_batch_device(swm_ctx_t *ctx, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                                    uint64_t n, uint32_t e, swm_read_result_t *d_out)
{
    int rc = check_map_args(ctx, stride, e);
    if (rc) return rc;
    if (n && (!d_reads || !d_lens || !d_out)) return SWM_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    return map_device_impl(ctx, d_reads, stride, d_lens, n, e, d_out);
}

extern "C" int swm_map_batch(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens, uint64_t n,
                             uint32_t e, swm_read_result_t *out)
{
    int rc = check_map_args(ctx, stride, e);
    if (rc) return rc;
    if (n == 0) return SWM_OK;
    if (!reads || !lens || !out) return SWM_EINVAL;
    for (uint64_t i = 0; i < n; i++)
        if ((uint64_t)lens[i] + 2 > stride) return SWM_EINVAL;   /* kseq stride rule (Q20) */
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    char *d_reads = nullptr;
    uint32_t *d_lens = nullptr;
    swm_read_result_t *d_out = nullptr;
    rc = SWM_OK;
    if (hipMalloc((void **)&d_reads, n * stride) != hipSuccess || hipMalloc((void **)&d_lens, n * 4) != hipSuccess ||
        hipMalloc((void **)&d_out, n * sizeof(swm_read_result_t)) != hipSuccess)
        rc = SWM_ENOMEM;
    if (rc == SWM_OK &&
        (hipMemcpyAsync(d_reads, reads, n * stride, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
         hipMemcpyAsync(d_lens, lens, n * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        rc = SWM_EDEVICE;
    if (rc == SWM_OK) rc = map_device_impl(ctx, d_reads, stride, d_lens, n, e, d_out);
    if (rc == SWM_OK && (hipMemcpyAsync(out, d_out, n * sizeof(swm_read_result_t), hipMemcpyDeviceToHost,
                                        ctx->stream) != hipSuccess ||
                         hipStreamSynchronize(ctx->stream) != hipSuccess))
        rc = SWM_EDEVICE;
    if (d_reads) (void)hipFree(d_reads);
    if (d_lens) (void)hipFree(d_lens);
    if (d_out) (void)hipFree(d_out);
    return rc;
}

extern "C" int swm_myers_device(swm_ctx_t *ctx, const uint32_t *d_read_words, const uint32_t *d_lens,
                                const uint32_t *d_cand_read, const uint32_t *d_win_start, const uint32_t *d_ref_len,
                                uint64_t n, uint32_t e, int32_t *d_err, int32_t *d_loc)
{
    if (!ctx || e > SWM_MAX_E) return SWM_EINVAL;
    if (n == 0) return SWM_OK;
    if (n >= 0xffffffffull || !d_read_words || !d_lens || !d_cand_read || !d_win_start || !d_ref_len || !d_err ||
        !d_loc)
        return SWM_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    k_myers_explicit<<<blocks_for(n, 256), 256, 0, ctx->stream>>>((uint32_t)n, d_read_words, d_lens, d_cand_read,
                                                                  d_win_start, d_ref_len, ctx->d_base_2bits, e,
                                                                  d_err, d_loc);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SWM_OK;
}

/* ------------------------------------------------------- device plumbing */

extern "C" int swm_device_count(int *n)
{
    if (!n) return SWM_EINVAL;
    *n = 0;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return SWM_OK;
    *n = c;
    return SWM_OK;
}

extern "C" int swm_dev_alloc(void **p, uint64_t bytes)
{
    if (!p) return SWM_EINVAL;
    *p = nullptr;
    return hipMalloc(p, bytes ? bytes : 1) == hipSuccess ? SWM_OK : SWM_ENOMEM;
}

extern "C" int swm_dev_free(void *p)
{
    return (!p || hipFree(p) == hipSuccess) ? SWM_OK : SWM_EDEVICE;
}

extern "C" int swm_copy_h2d(void *dst, const void *src, uint64_t bytes)
{
    if (bytes && (!dst || !src)) return SWM_EINVAL;
    return (!bytes || hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess) ? SWM_OK : SWM_EDEVICE;
}

extern "C" int swm_copy_d2h(void *dst, const void *src, uint64_t bytes)
{
    if (bytes && (!dst || !src)) return SWM_EINVAL;
    return (!bytes || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess) ? SWM_OK : SWM_EDEVICE;
}

extern "C" int swm_dev_memset(void *dst, int value, uint64_t bytes)
{
    if (bytes && !dst) return SWM_EINVAL;
    return (!bytes || hipMemset(dst, value, bytes) == hipSuccess) ? SWM_OK : SWM_EDEVICE;
}
