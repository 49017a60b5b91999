/*
 * swm_index.hip -- 12-mer hash index build on the GPU, byte-identical to
 * serial_index.c under HOST_BUILD (construct_hash_index, :359-445).
 *
 *   k_encode   one thread per 32-base word: bit planes (.base_2bits), N mask
 *              (.base_3bits) and the build-only 2-bit seed words
 *              (get_seed/get_base, serial_index.c:146-235), including the
 *              final flush that zeroes the last word when the genome length
 *              is a multiple of the word size (:423-426)
 *   k_pack_hist  every sort pass's digit counts from the genome words: the
 *              sampled starts p = 0, 4, 8, ... <= N-k and whether their
 *              k-mer has an N (the histogram's starts, :253-269)
 *   k_pack_pass0 each start as one 64-bit key code | residual << 2k |
 *              (p/4) << (2k+12) -- the k-mer code from the 2-bit seed words,
 *              the N test from the mask, the 12-bit residual
 *              (get_seed_from_enc_bits :280-331) from the planes, the last
 *              starts (the tail below) from the sequence itself -- written
 *              straight to its place after the first sort pass, the starts
 *              with an N after all valid ones
 *   sort       stable LSD radix sort on the code bits only (3 passes of 8
 *              bits for k=12, k_rs_*): the counting sort's bucket order
 *              with positions ascending (:343-352); its last pass writes
 *              .positions and .pos_lseeds in bucket order and the bucket
 *              starts (store_positions' offsets :337-340) at each change of
 *              code, k_rs_fix those at the first key of each tile's run
 *   k_counts   .hash_count and the END offsets from the starts
 *   k_tail     one wave replays the scatter for the last starts, whose
 *              scatter reads the zeroed words, so that slots, overflow and
 *              the END offsets of .hash_offset (Q19) match the serial build
 * Genomes whose key would not fit 64 bits (more than 2^(52-2k) sampled
 * starts) sort code << 32 | p on the code bits instead and take the residual
 * from the planes afterwards.
 */
#include "swm_internal.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>

namespace {

/* toupper (serial_index.c:152), then A, C, G, T -> 0..3, anything else -1;
 * branch-free (a switch here compiled to a branch per byte). */
__device__ __forceinline__ int base_code(char raw)
{
    const uint32_t c = (uint8_t)raw;
    const uint32_t u = (c - (uint32_t)'a' < 26u) ? c - 32u : c;
    const uint32_t code = ((u >> 1) ^ (u >> 2)) & 3u;      /* A 0, C 1, G 2, T 3 */
    const bool valid = (u == 'A') | (u == 'C') | (u == 'G') | (u == 'T');
    return valid ? (int)code : -1;
}

/* Inclusive scan of one u32 per thread over the workgroup (NT threads). */
template <int NT>
__device__ __forceinline__ uint32_t block_incl_scan_u32(uint32_t v, uint32_t *s_w, uint32_t *total)
{
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, d, 64);
        if (lane >= (uint32_t)d) v += u;
    }
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
        const uint32_t x = s_w[i];
        if ((uint32_t)i < w) off += x;
        tot += x;
    }
    *total = tot;
    return v + off;
}

/* Also clears zero[0 .. nzero) (the sort's digit counts and tile counters). */
__global__ void __launch_bounds__(256) k_encode(const char *__restrict__ seq, uint64_t n, uint64_t nwords,
                                                uint32_t *__restrict__ b2, uint32_t *__restrict__ b3,
                                                uint32_t *__restrict__ s2, uint32_t *__restrict__ zero,
                                                uint32_t nzero)
{
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = w; i < nzero; i += (uint64_t)gridDim.x * blockDim.x) zero[i] = 0;
    if (w >= nwords) return;
    uint32_t lo = 0, hi = 0, nm = 0, s0 = 0, s1 = 0;
    const uint64_t p0 = w * 32;
    if (p0 + 32 <= n) {
        const uint4 *v = reinterpret_cast<const uint4 *>(seq + p0);
        uint4 a = v[0], b = v[1];
        uint32_t words[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int j = 0; j < 32; j++) {
            int c = base_code((char)((words[j >> 2] >> ((j & 3) * 8)) & 0xffu));
            uint32_t cc = c < 0 ? 0u : (uint32_t)c;
            lo |= (cc & 1u) << j;
            hi |= (cc >> 1) << j;
            nm |= (c < 0 ? 1u : 0u) << j;
            if (j < 16) s0 |= cc << (2 * j);
            else s1 |= cc << (2 * (j - 16));
        }
    } else {
        for (int j = 0; j < 32 && p0 + j < n; j++) {
            int c = base_code(seq[p0 + j]);
            uint32_t cc = c < 0 ? 0u : (uint32_t)c;
            lo |= (cc & 1u) << j;
            hi |= (cc >> 1) << j;
            nm |= (c < 0 ? 1u : 0u) << j;
            if (j < 16) s0 |= cc << (2 * j);
            else s1 |= cc << (2 * (j - 16));
        }
    }
    /* final flush of the empty accumulators (serial_index.c:423-426) */
    if (n % 32 == 0 && w == nwords - 1) { lo = 0; hi = 0; nm = 0; }
    if (n % 16 == 0) {
        uint64_t last_s = n / 16 - 1;
        if (2 * w == last_s) s0 = 0;
        if (2 * w + 1 == last_s) s1 = 0;
    }
    b2[2 * w] = lo;
    b2[2 * w + 1] = hi;
    b3[w] = nm;
    s2[2 * w] = s0;
    s2[2 * w + 1] = s1;
}

/* get_seed_from_enc_bits residual (serial_index.c:290-315) from the planes
 * as stored. */
__device__ __forceinline__ uint32_t lseed_at(const uint32_t *__restrict__ b2, uint64_t pos, uint32_t k,
                                             uint64_t pair0 = 0)
{
    uint64_t s_pos = pos >= 3 ? pos - 3 : 0;
    uint64_t wi = (s_pos >> 5) - pair0;
    uint32_t sh = (uint32_t)s_pos & 31u;
    uint32_t fl = __builtin_amdgcn_alignbit(b2[2 * wi + 2], b2[2 * wi], sh) & 7u;
    uint32_t fh = __builtin_amdgcn_alignbit(b2[2 * wi + 3], b2[2 * wi + 1], sh) & 7u;
    s_pos = pos + k;
    wi = (s_pos >> 5) - pair0;
    sh = (uint32_t)s_pos & 31u;
    uint32_t bl = __builtin_amdgcn_alignbit(b2[2 * wi + 2], b2[2 * wi], sh) & 7u;
    uint32_t bh = __builtin_amdgcn_alignbit(b2[2 * wi + 3], b2[2 * wi + 1], sh) & 7u;
    fl |= bl << 3;
    fh |= bh << 3;
    return fl | (fh << 6);
}

/* Large genomes: one thread per sampled start p = 0, 4, 8, ... <= N-k, the
 * key code << 32 | p with the k-mer code and N-freeness of the histogram
 * (:253-269), 4^k for a start with an N (sorted after every valid one). */
__global__ void __launch_bounds__(256) k_keys(const char *__restrict__ seq, uint64_t nsamp, uint32_t k,
                                              unsigned long long *__restrict__ keys)
{
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nsamp) return;
    const uint64_t p = j * SWM_SEED_STEP;
    /* starts are multiples of 4, so the k <= 14 bases come in four aligned
     * 32-bit loads (the device copy of the sequence has 64 spare bytes) */
    const uint32_t *w = reinterpret_cast<const uint32_t *>(seq + p);
    const uint32_t wv[4] = {w[0], w[1], w[2], w[3]};
    uint32_t code = 0;
    bool valid = true;
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) {
        if (i < k) {
            const int c = base_code((char)((wv[i >> 2] >> (8 * (i & 3))) & 0xffu));
            if (c < 0) valid = false;
            code |= (uint32_t)(c < 0 ? 0 : c) << (2 * i);
        }
    }
    keys[j] = ((unsigned long long)(valid ? code : (1u << (2 * k))) << 32) | (uint32_t)p;
}

/* The histogram's code and N test of start p: from the seed words and the
 * mask for p < t0, from the sequence for the last starts (whose words the
 * final flush may have zeroed). */
__device__ __forceinline__ bool start_code(const char *__restrict__ seq, const uint32_t *__restrict__ s2,
                                           const uint32_t *__restrict__ b3, uint64_t p, uint64_t t0, uint32_t k,
                                           uint32_t *code, uint64_t s2o = 0, uint64_t b3o = 0)
{
    if (p < t0) {
        const uint64_t sw = (p >> 4) - s2o, mw = (p >> 5) - b3o;
        const uint32_t smask = (uint32_t)((1ull << (2 * k)) - 1), fmask = (1u << k) - 1u;
        *code = __builtin_amdgcn_alignbit(s2[sw + 1], s2[sw], (uint32_t)(2 * (p & 15))) & smask;
        return (__builtin_amdgcn_alignbit(b3[mw + 1], b3[mw], (uint32_t)(p & 31)) & fmask) == 0;
    }
    uint32_t c = 0;
    bool valid = true;
    for (uint32_t i = 0; i < k; i++) {
        const int b = base_code(seq[p + i]);
        if (b < 0) valid = false;
        c |= (uint32_t)(b < 0 ? 0 : b) << (2 * i);
    }
    *code = c;
    return valid;
}

/* The genome words that the sampled starts j in [tb, te) (te - tb <=
 * RS_TILE) read, staged in LDS with coalesced loads: seed words s2[tb/4 ..
 * (te-1)/4 + 1], mask words b3[tb/8 .. (te-1)/8 + 1] and plane pairs from
 * the residual's first word ((4 tb - 3) / 32) to its last ((4 (te-1) + k) /
 * 32 + 1).  Every start's code, N test and residual then read LDS. */
#define STG_S2 0
#define STG_B3 1040
#define STG_B2 1568
#define STG_WORDS 2624

struct Stage {
    uint64_t s2o, b3o, b2o;                /* first staged seed word, mask word, plane pair */
};

__device__ __forceinline__ Stage stage_tile(const uint32_t *__restrict__ s2, const uint32_t *__restrict__ b3,
                                            const uint32_t *__restrict__ b2, uint64_t tb, uint64_t te, uint32_t k,
                                            uint32_t *lds, uint32_t tid, uint32_t nt)
{
    Stage g;
    const uint64_t p0 = tb * SWM_SEED_STEP, p1 = (te - 1) * SWM_SEED_STEP;
    g.s2o = tb >> 2;
    g.b3o = tb >> 3;
    g.b2o = (p0 >= 3 ? p0 - 3 : 0) >> 5;
    const uint32_t ns2 = (uint32_t)(((te - 1) >> 2) + 2 - g.s2o);         /* <= 1025 */
    const uint32_t nb3 = (uint32_t)(((te - 1) >> 3) + 2 - g.b3o);         /* <= 513 */
    const uint32_t nb2 = (uint32_t)(2 * (((p1 + k) >> 5) + 2 - g.b2o));   /* <= 1034 */
    for (uint32_t i = tid; i < ns2; i += nt) lds[STG_S2 + i] = s2[g.s2o + i];
    for (uint32_t i = tid; i < nb3; i += nt) lds[STG_B3 + i] = b3[g.b3o + i];
    for (uint32_t i = tid; i < nb2; i += nt) lds[STG_B2 + i] = b2[2 * g.b2o + i];
    return g;
}

/* Counts and END offsets from the starts. */
__global__ void __launch_bounds__(256) k_counts(uint64_t nb, const uint32_t *__restrict__ start,
                                                uint32_t *__restrict__ count, uint32_t *__restrict__ end)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nb) {
        const uint32_t a = start[c], b = start[c + 1];
        count[c] = b - a;
        end[c] = b;
    } else if (c == nb) {
        end[nb] = start[nb];
    }
}

/* Bucket counts from the sorted keys: the thread at the end of each run of
 * one valid code gallops back to the run's start (runs average a few
 * entries; long ones cost log steps). */
__global__ void __launch_bounds__(256) k_runs(uint64_t nsamp, uint32_t nb, const unsigned long long *__restrict__ kv,
                                              uint32_t *__restrict__ count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsamp) return;
    auto key = [&](uint64_t x) { return (uint32_t)(kv[x] >> 32); };
    const uint32_t c = key(i);
    if (c >= nb || (i + 1 < nsamp && key(i + 1) == c)) return;
    uint64_t lo = i, step = 1;             /* key(lo) == c */
    while (step <= lo && key(lo - step) == c) { lo -= step; step <<= 1; }
    uint64_t a = step <= lo ? lo - step : 0;   /* key(a) != c or a == 0 */
    if (key(a) == c) lo = a;
    else {
        while (lo - a > 1) {               /* key(a) != c, key(lo) == c */
            const uint64_t mid = a + (lo - a) / 2;
            if (key(mid) == c) lo = mid; else a = mid;
        }
    }
    count[c] = (uint32_t)(i + 1 - lo);
}

/* positions and residuals in bucket order (the residual from the planes) */
__global__ void __launch_bounds__(256) k_split(const uint64_t *__restrict__ mp, const unsigned long long *__restrict__ v,
                                               const uint32_t *__restrict__ b2, uint32_t k,
                                               uint32_t *__restrict__ positions, uint16_t *__restrict__ lseeds)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *mp) return;
    const uint32_t p = (uint32_t)v[i];
    positions[i] = p;
    lseeds[i] = (uint16_t)lseed_at(b2, p, k);
}

/* M = start[nb] (the pair path's total). */
__global__ void k_m_from_start(const uint32_t *__restrict__ start, uint64_t nb, uint64_t *__restrict__ m)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *m = start[nb];
}

__global__ void k_end_offsets(uint64_t nb, const uint32_t *__restrict__ start, uint32_t *__restrict__ end)
{
    uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nb) end[c] = start[c + 1];
    if (c == nb) end[nb] = start[nb];
}

#define TAIL_MAX 64

/* Replay of store_positions (serial_index.c:343-352) for the starts p >= t0,
 * whose scatter code and N check read the (possibly zeroed) last words; all
 * earlier starts read the same bits as the histogram.  One wave, lane i on
 * tail start i (at most (64+k)/4 + 1 of them): the codes, ranks and slots
 * are computed lane-parallel, the three store phases are ordered by fences. */
__global__ void __launch_bounds__(64) k_tail(const char *__restrict__ seq, uint64_t n, uint32_t k, uint64_t t0,
                                             const uint32_t *__restrict__ b2, const uint32_t *__restrict__ b3,
                                             const uint32_t *__restrict__ s2, const uint32_t *__restrict__ count,
                                             const uint32_t *__restrict__ start, const uint64_t *__restrict__ mp,
                                             uint32_t *__restrict__ positions, uint16_t *__restrict__ lseeds,
                                             uint32_t *__restrict__ end)
{
    if (blockIdx.x != 0) return;
    const uint32_t i = threadIdx.x;
    const uint64_t M = *mp;
    const uint64_t nt64 = (t0 + k <= n) ? (n - k - t0) / SWM_SEED_STEP + 1 : 0;
    const uint32_t nt = (uint32_t)(nt64 < TAIL_MAX ? nt64 : TAIL_MAX);
    const bool live = i < nt;
    const uint64_t p = t0 + (uint64_t)i * SWM_SEED_STEP;
    const uint32_t smask = (uint32_t)((1ull << (2 * k)) - 1), fmask = (uint32_t)((1ull << k) - 1);
    uint32_t tcode = 0, scode = 0;
    bool tvalid = false, svalid = false;
    if (live) {
        tvalid = true;
        for (uint32_t q = 0; q < k; q++) {
            const int c = base_code(seq[p + q]);
            if (c < 0) tvalid = false;
            tcode |= (uint32_t)(c < 0 ? 0 : c) << (2 * q);
        }
        scode = __builtin_amdgcn_alignbit(s2[p / 16 + 1], s2[p / 16], (uint32_t)((p % 16) * 2)) & smask;
        svalid = (__builtin_amdgcn_alignbit(b3[p / 32 + 1], b3[p / 32], (uint32_t)(p % 32)) & fmask) == 0;
    }
    /* per lane: tail starts of its true code (all / before it), scatter
     * starts of its scatter code before it, and of its true code in all */
    uint32_t t_same_t = 0, t_rank = 0, s_prior = 0, s_writes_t = 0, t_same_s = 0, s_writes_s = 0;
    const uint64_t tvm = __ballot(tvalid), svm = __ballot(svalid);
    for (uint32_t q = 0; q < nt; q++) {
        const uint32_t tc = (uint32_t)__builtin_amdgcn_readlane((int)tcode, (int)q);
        const uint32_t sc = (uint32_t)__builtin_amdgcn_readlane((int)scode, (int)q);
        const bool tv = (tvm >> q) & 1ull, sv = (svm >> q) & 1ull;
        if (tv && tc == tcode) { t_same_t++; if (q < i) t_rank++; }
        if (tv && tc == scode) t_same_s++;
        if (sv && sc == scode) { s_writes_s++; if (q < i) s_prior++; }
        if (sv && sc == tcode) s_writes_t++;
    }
    const uint32_t lseed = live ? lseed_at(b2, p, k) : 0u;
    /* 1) clear the slots the bucket sort gave to tail starts (the serial
     *    scatter writes them only through the replay below) */
    if (live && tvalid) {
        const uint64_t slot = (uint64_t)start[tcode] + (count[tcode] - t_same_t) + t_rank;
        if (slot < M) { positions[slot] = 0; lseeds[slot] = 0; }
    }
    __threadfence();
    /* 2) the serial scatter of the tail: in start order, later writes win
     *    (two starts can meet in one slot when a bucket overflows), so the
     *    lanes store one after the other */
    for (uint32_t q = 0; q < nt; q++) {
        if (i == q && svalid) {
            const uint64_t slot = (uint64_t)start[scode] + (count[scode] - t_same_s) + s_prior;
            if (slot < M) { positions[slot] = (uint32_t)p; lseeds[slot] = (uint16_t)lseed; }
        }
        __threadfence();
    }
    /* 3) END offsets of every bucket the tail touches */
    if (live && tvalid) end[tcode] = start[tcode] + (count[tcode] - t_same_t) + s_writes_t;
    if (live && svalid) end[scode] = start[scode] + (count[scode] - t_same_s) + s_writes_s;
}

/* ------------------------------------------------------------- scans */

#define SC_NT 256
#define SC_PER 16
#define SC_TILE (SC_NT * SC_PER)

/* Exclusive scan of u32 (sums fit u32: they are index offsets) in two
 * launches: tile sums, then each tile adds the sums before it and scans in
 * LDS (coalesced in and out). */
__global__ void __launch_bounds__(SC_NT) k_sc_reduce(const uint32_t *__restrict__ in, uint64_t n,
                                                     uint32_t *__restrict__ sums)
{
    __shared__ uint32_t s_w[SC_NT / 64];
    const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < SC_PER; q++) {
        const uint64_t g = base + (uint64_t)q * SC_NT + threadIdx.x;
        v += g < n ? in[g] : 0u;
    }
    uint32_t tot;
    (void)block_incl_scan_u32<SC_NT>(v, s_w, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SC_NT) k_sc_down(const uint32_t *__restrict__ in, uint64_t n,
                                                   const uint32_t *__restrict__ sums, uint32_t *__restrict__ out)
{
    __shared__ uint32_t s_w[SC_NT / 64];
    __shared__ uint32_t s_v[SC_TILE + SC_TILE / 16];
    const uint64_t base = (uint64_t)blockIdx.x * SC_TILE;
    uint32_t pre = 0;
    for (uint32_t b = threadIdx.x; b < blockIdx.x; b += SC_NT) pre += sums[b];
    uint32_t tile_off;
    (void)block_incl_scan_u32<SC_NT>(pre, s_w, &tile_off);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SC_PER; q++) {
        const uint32_t e = (uint32_t)q * SC_NT + threadIdx.x;
        const uint64_t g = base + e;
        s_v[e + (e >> 4)] = g < n ? in[g] : 0u;
    }
    __syncthreads();
    uint32_t v[SC_PER], t = 0;
#pragma unroll
    for (int q = 0; q < SC_PER; q++) {
        const uint32_t e = threadIdx.x * SC_PER + (uint32_t)q;
        v[q] = s_v[e + (e >> 4)];
        t += v[q];
    }
    uint32_t tot;
    uint32_t run = block_incl_scan_u32<SC_NT>(t, s_w, &tot) - t + tile_off;
#pragma unroll
    for (int q = 0; q < SC_PER; q++) {
        const uint32_t e = threadIdx.x * SC_PER + (uint32_t)q;
        s_v[e + (e >> 4)] = run;
        run += v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SC_PER; q++) {
        const uint32_t e = (uint32_t)q * SC_NT + threadIdx.x;
        const uint64_t g = base + e;
        if (g < n) out[g] = s_v[e + (e >> 4)];
    }
}

/* ------------------------------------------------------ radix sort */

/* Stable LSD radix sort of 64-bit keys on bits [lo, hi), 8 bits a pass.
 * The digit starts of every pass come from one histogram of the input
 * (k_rs_hist, or the kernel that writes the keys).  Each pass (k_rs_pass)
 * takes tiles of 4096 keys in the order their workgroups start (a counter,
 * so every tile's predecessors are running): it counts the tile's digits
 * and publishes the counts at once, finds its global offsets by a decoupled
 * look-back over the tiles before it (RS_LB statuses read per step, summed
 * up to the first that includes every tile before it), ranks the keys in
 * the tile and scatters them through LDS so that each digit's run of the
 * tile is written contiguously.  Measured on 64M keys (one pass): 438 us
 * with an 8-wide look-back step, 465 with 16, 485 with 24; 265-300 us with
 * no waiting at all; 490 us for the wait-free form (a count pass, a scan of
 * the tiles' counts, a scatter pass); two-level look-back over groups of
 * 16-64 tiles 495-520 us.
 *
 * Stability: wave w of a tile holds keys w*1024 + r*64 + lane (r = 0..15),
 * the keys' own order; a wave ranks its rounds in order against a private
 * per-digit counter (lanes of one digit by a match over the digit's bits),
 * then the waves' counters are offset in wave order. */
#define RS_NT 256
#define RS_PER 16
#define RS_WAVES (RS_NT / 64)
#define RS_TILE (RS_NT * RS_PER)
#define RS_AGG 0x40000000u       /* status: the tile's own count */
#define RS_PRE 0x80000000u       /* status: the count of this and every earlier tile */
#define RS_VAL 0x3fffffffu       /* counts below 2^30 (fewer than 2^30 sampled starts) */
#define RS_SPIN (1u << 24)       /* a look-back gives up after this many waits (and flags it) */
#define RS_LB 8                  /* predecessors a look-back step reads at once */

/* Thread d of tile `tile` in a pass: publishes the tile's count of digit d
 * and returns the count of digit d over every earlier tile (decoupled
 * look-back).  RS_LB predecessors are read at once, nearest first: their
 * counts are added up to the first inclusive one; an unpublished one stops
 * the window, which is read again from there.  Tiles start in the order of
 * the tile counter, so every predecessor is running or done. */
__device__ __forceinline__ uint32_t rs_lookback(uint32_t *__restrict__ status, uint32_t tile, uint32_t d,
                                                uint32_t tot, uint32_t *__restrict__ err)
{
    uint32_t *my = status + (uint64_t)tile * 256 + d;
    __hip_atomic_store(my, (tile == 0 ? RS_PRE : RS_AGG) | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t pre = 0;
    if (tile == 0) return 0;
    uint32_t spins = 0;
    int64_t j = (int64_t)tile - 1;
    for (;;) {
        uint32_t x[RS_LB];
#pragma unroll
        for (int q = 0; q < RS_LB; q++)
            x[q] = j - q >= 0 ? __hip_atomic_load(status + (uint64_t)(j - q) * 256 + d, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : RS_PRE;
        bool done = false;
        int used = RS_LB;
#pragma unroll
        for (int q = 0; q < RS_LB; q++) {
            if (done || used != RS_LB) continue;
            if ((x[q] & (RS_AGG | RS_PRE)) == 0) { used = q; continue; }
            pre += x[q] & RS_VAL;
            if (x[q] & RS_PRE) done = true;
        }
        if (done) break;
        j -= used;
        if (used < RS_LB) {
            if (++spins > RS_SPIN) { atomicOr(err, 1u); break; }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __hip_atomic_store(my, RS_PRE | ((pre + tot) & RS_VAL), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return pre;
}

/* The passes of a sort: pass p orders on bits [sh[p], sh[p] + bits[p]). */
struct PassPlan {
    uint32_t n;
    uint32_t sh[4], bits[4];
};

__device__ __forceinline__ uint32_t pass_digit(unsigned long long x, const PassPlan &pl, uint32_t p)
{
    return (uint32_t)(x >> pl.sh[p]) & ((1u << pl.bits[p]) - 1u);
}

__global__ void __launch_bounds__(256) k_rs_hist(const unsigned long long *__restrict__ keys, uint64_t n,
                                                 PassPlan pl, uint32_t *__restrict__ hist)
{
    __shared__ uint32_t h[4][4 * 256];
    uint32_t *hw = h[threadIdx.x >> 6];
    for (uint32_t i = threadIdx.x; i < 4 * 4 * 256; i += blockDim.x) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride) {
        unsigned long long x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const uint64_t i = i0 + (uint64_t)u * stride;
            x[u] = keys[i < n ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (i0 + (uint64_t)u * stride >= n) continue;
            for (uint32_t p = 0; p < pl.n; p++) atomicAdd(&hw[p * 256 + pass_digit(x[u], pl, p)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < pl.n * 256; i += blockDim.x) {
        const uint32_t c = h[0][i] + h[1][i] + h[2][i] + h[3][i];
        if (c) atomicAdd(&hist[i], c);
    }
}

/* Digit starts of every pass: exclusive scans of the 256 counts; total0
 * (if given) gets pass 0's total. */
__global__ void __launch_bounds__(256) k_rs_bases(const uint32_t *__restrict__ hist, uint32_t npass,
                                                  uint32_t *__restrict__ base, uint64_t *__restrict__ total0)
{
    __shared__ uint32_t s_w[4];
    for (uint32_t p = 0; p < npass; p++) {
        const uint32_t c = hist[p * 256 + threadIdx.x];
        uint32_t tot;
        const uint32_t incl = block_incl_scan_u32<256>(c, s_w, &tot);
        base[p * 256 + threadIdx.x] = incl - c;
        if (p == 0 && total0 && threadIdx.x == 0) *total0 = tot;
        __syncthreads();
    }
}

/* What the last pass of the index sort writes instead of keys: for the
 * sorted key at index i < M (key = code | residual << 2k | j << (2k + 12)),
 * positions[i] = j * SWM_SEED_STEP and lseeds[i] = residual, and the bucket
 * starts of the code boundaries inside a tile's digit run (and of the codes
 * above the last valid key).  The boundaries at the first key of each run
 * are left to k_rs_fix, for which each tile records per digit the code of
 * its run's first key (rf) and the code of its last key plus one (rl, 0 for
 * an empty run). */
struct SortFinish {
    const uint64_t *m;
    uint32_t k;
    uint64_t nb;
    uint32_t *start, *positions;
    uint16_t *lseeds;
    uint32_t *rf, *rl;
};

template <bool FIN>
__global__ void __launch_bounds__(RS_NT) k_rs_pass(const unsigned long long *__restrict__ in,
                                                   unsigned long long *__restrict__ out, uint64_t n, uint32_t shift,
                                                   uint32_t bits, const uint32_t *__restrict__ base,
                                                   uint32_t *__restrict__ status, uint32_t *__restrict__ tile_ctr,
                                                   uint32_t *__restrict__ err, SortFinish fin)
{
    __shared__ unsigned long long s_k[RS_TILE];
    __shared__ uint32_t s_wc[RS_WAVES * 256];
    __shared__ uint32_t s_dstart[256], s_gofs[256], s_cnt[256], s_w[RS_WAVES];
    __shared__ uint32_t s_tile;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (uint32_t i = tid; i < RS_WAVES * 256; i += RS_NT) s_wc[i] = 0;
    if (tid < 256) s_cnt[tid] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t t0 = (uint64_t)tile * RS_TILE + (uint64_t)w * (RS_TILE / RS_WAVES);
    unsigned long long key[RS_PER];
    uint32_t rank[RS_PER];
#pragma unroll
    for (int r = 0; r < RS_PER; r++) {
        const uint64_t g = t0 + (uint64_t)r * 64 + lane;
        key[r] = in[g < n ? g : 0];
    }
    /* the tile's digit counts first, published at once: the tiles after it
     * find them while this one ranks its keys */
#pragma unroll
    for (int r = 0; r < RS_PER; r++)
        if (t0 + (uint64_t)r * 64 + lane < n) atomicAdd(&s_cnt[(uint32_t)(key[r] >> shift) & mask], 1u);
    __syncthreads();
    if (tid < 256) {
        /* before the ranking, so that this tile's inclusive count is out
         * as early as possible */
        s_gofs[tid] = base[tid] + rs_lookback(status, tile, tid, s_cnt[tid], err);
    }
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t *wc = s_wc + w * 256;
#pragma unroll
    for (int r = 0; r < RS_PER; r++) {
        const bool v = t0 + (uint64_t)r * 64 + lane < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        uint64_t peers = __ballot(v);
        for (uint32_t b = 0; b < bits; b++) {
            const uint64_t bb = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t before = wc[d];
        rank[r] = before + (uint32_t)__popcll(peers & below);
        if (v && (peers & below) == 0) wc[d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        /* thread d: the waves' offsets within digit d and the digit's start
         * within the tile */
        const uint32_t d = tid;
        uint32_t tot = 0;
        if (d < 256) {
#pragma unroll
            for (int q = 0; q < RS_WAVES; q++) {
                const uint32_t c = s_wc[q * 256 + d];
                s_wc[q * 256 + d] = tot;
                tot += c;
            }
        }
        uint32_t all;
        const uint32_t ds = block_incl_scan_u32<RS_NT>(tot, s_w, &all) - tot;
        if (d < 256) s_dstart[d] = ds;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_PER; r++) {
        const bool v = t0 + (uint64_t)r * 64 + lane < n;
        const uint32_t d = (uint32_t)(key[r] >> shift) & mask;
        if (v) s_k[s_dstart[d] + s_wc[w * 256 + d] + rank[r]] = key[r];
    }
    __syncthreads();
    const uint64_t tb = (uint64_t)tile * RS_TILE;
    const uint32_t nv = (uint32_t)(n - tb < RS_TILE ? n - tb : RS_TILE);
    if (!FIN) {
        for (uint32_t i = tid; i < nv; i += RS_NT) {
            const unsigned long long x = s_k[i];
            const uint32_t d = (uint32_t)(x >> shift) & mask;
            out[(uint64_t)s_gofs[d] + (i - s_dstart[d])] = x;
        }
        return;
    }
    const uint64_t m = *fin.m;
    const uint32_t cb = 2 * fin.k;
    const unsigned long long cm = (1ull << cb) - 1ull;
    if (tid < 256) {
        const uint32_t c = s_cnt[tid];
        uint32_t f = 0, l = 0;
        if (c) {
            f = (uint32_t)(s_k[s_dstart[tid]] & cm);
            l = (uint32_t)(s_k[s_dstart[tid] + c - 1] & cm) + 1u;
        }
        fin.rf[(uint64_t)tile * 256 + tid] = f;
        fin.rl[(uint64_t)tile * 256 + tid] = l;
    }
    for (uint32_t i = tid; i < nv; i += RS_NT) {
        const unsigned long long x = s_k[i];
        const uint32_t d = (uint32_t)(x >> shift) & mask;
        const uint64_t gi = (uint64_t)s_gofs[d] + (i - s_dstart[d]);
        if (gi >= m) continue;            /* the starts with an N, after every valid one */
        fin.positions[gi] = (uint32_t)((x >> (cb + 12)) * SWM_SEED_STEP);
        fin.lseeds[gi] = (uint16_t)((x >> cb) & 0xfffu);
        const uint64_t code = x & cm;
        if (i > s_dstart[d]) {
            /* the key before it in the output is the one before it here */
            const uint64_t prev = s_k[i - 1] & cm;
            for (uint64_t c = prev + 1; c <= code; c++) fin.start[c] = (uint32_t)gi;
        }
        if (gi + 1 == m)
            for (uint64_t c = code + 1; c <= fin.nb; c++) fin.start[c] = (uint32_t)m;
    }
}

/* The packed index's digit counts before its keys exist: each sampled
 * start's key has the start's code in its low bits, or all ones for a start
 * with an N; these are counted for every pass of the sort, pass 0 over the
 * valid starts only (k_pack_pass0 puts the starts with an N after all of
 * them).  Workgroups take tiles of RS_TILE starts grid-stride, their words
 * staged in LDS.  Also clears every pass's look-back status. */
__global__ void __launch_bounds__(256) k_pack_hist(const char *__restrict__ seq, uint64_t nsamp, uint32_t k,
                                                   uint64_t t0, const uint32_t *__restrict__ s2,
                                                   const uint32_t *__restrict__ b3, const uint32_t *__restrict__ b2,
                                                   PassPlan pl, uint32_t *__restrict__ hist,
                                                   uint32_t *__restrict__ status, uint64_t status_n)
{
    __shared__ uint32_t h[4][4 * 256];     /* one histogram per wave: waves do not contend */
    __shared__ uint32_t stg[STG_WORDS];
    const uint32_t tid = threadIdx.x;
    uint32_t *hw = h[tid >> 6];
    for (uint32_t i = tid; i < 4 * 4 * 256; i += 256) (&h[0][0])[i] = 0;
    const uint64_t gid = (uint64_t)blockIdx.x * 256 + tid, gs = (uint64_t)gridDim.x * 256;
    for (uint64_t i = gid; i < status_n; i += gs) status[i] = 0;
    const uint32_t cm = (uint32_t)((1ull << (2 * k)) - 1);
    const uint64_t ntiles = (nsamp + RS_TILE - 1) / RS_TILE;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t tb = tile * RS_TILE, te = min(tb + RS_TILE, nsamp);
        __syncthreads();                   /* the previous tile's words are read */
        const Stage g = stage_tile(s2, b3, b2, tb, te, k, stg, tid, 256);
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < RS_TILE / 256; r++) {
            const uint64_t j = tb + (uint64_t)r * 256 + tid;
            if (j >= te) break;
            uint32_t code;
            const bool ok = start_code(seq, stg + STG_S2, stg + STG_B3, j * SWM_SEED_STEP, t0, k, &code, g.s2o,
                                       g.b3o);
            const uint32_t low = ok ? code : cm;
            if (ok) atomicAdd(&hw[pass_digit(low, pl, 0)], 1u);
            for (uint32_t p = 1; p < pl.n; p++) atomicAdd(&hw[p * 256 + pass_digit(low, pl, p)], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < pl.n * 256; i += 256) {
        const uint32_t c = h[0][i] + h[1][i] + h[2][i] + h[3][i];
        if (c) atomicAdd(&hist[i], c);
    }
}

/* The packed index's keys and the first pass of their sort in one kernel.
 * A tile is RS_TILE consecutive sampled starts j (tiles taken in workgroup
 * start order, as in k_rs_pass).  Each start's key -- code | residual << 2k
 * | j << (2k + 12), or all-ones code | j << (2k + 12) for a start with an N
 * -- is built from the genome words, ranked in the tile (valid starts by
 * their pass-0 digit, starts with an N as one more class after them) and
 * written where the stable pass would put the keys listed in genome order,
 * valid ones first: a valid key at base[d] + (digit d in earlier tiles, by
 * look-back) + its rank, a start with an N at M + (starts with an N in
 * earlier tiles) + its rank. */
__global__ void __launch_bounds__(RS_NT) k_pack_pass0(const char *__restrict__ seq, uint64_t nsamp, uint32_t k,
                                                      uint64_t t0s, const uint32_t *__restrict__ s2,
                                                      const uint32_t *__restrict__ b3, const uint32_t *__restrict__ b2,
                                                      uint32_t bits, const uint32_t *__restrict__ base,
                                                      const uint64_t *__restrict__ mp, uint32_t *__restrict__ status,
                                                      uint32_t *__restrict__ tile_ctr, uint32_t *__restrict__ err,
                                                      unsigned long long *__restrict__ out)
{
    constexpr uint32_t NC = 257;           /* classes: 256 digits, then the starts with an N */
    __shared__ unsigned long long s_k[RS_TILE];
    __shared__ uint32_t s_wc[RS_WAVES * NC];
    __shared__ uint32_t s_dstart[256], s_gofs[256], s_cnt[NC], s_w[RS_WAVES], s_w2[RS_WAVES];
    __shared__ uint32_t s_tile, s_nval;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    const uint32_t mask = (1u << bits) - 1u;
    if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
    for (uint32_t i = tid; i < RS_WAVES * NC; i += RS_NT) s_wc[i] = 0;
    for (uint32_t i = tid; i < NC; i += RS_NT) s_cnt[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t jw = (uint64_t)tile * RS_TILE + (uint64_t)w * (RS_TILE / RS_WAVES);
    const uint32_t cb = 2 * k;
    const unsigned long long cm = (1ull << cb) - 1ull;
    /* the tile's genome words in the key buffer (free until the keys are
     * placed, several barriers later) */
    uint32_t *stg = reinterpret_cast<uint32_t *>(s_k);
    const uint64_t tb0 = (uint64_t)tile * RS_TILE;
    const Stage g = stage_tile(s2, b3, b2, tb0, min(tb0 + RS_TILE, nsamp), k, stg, tid, RS_NT);
    __syncthreads();
    unsigned long long key[RS_PER];
    uint32_t rank[RS_PER];
    uint32_t nflag = 0;                    /* bit r: row r's start has an N */
#pragma unroll
    for (int r = 0; r < RS_PER; r++) {
        const uint64_t j = jw + (uint64_t)r * 64 + lane;
        uint32_t code = 0;
        const bool ok = j < nsamp && start_code(seq, stg + STG_S2, stg + STG_B3, j * SWM_SEED_STEP, t0s, k, &code,
                                                g.s2o, g.b3o);
        if (ok) {
            key[r] = (unsigned long long)code |
                     ((unsigned long long)lseed_at(stg + STG_B2, j * SWM_SEED_STEP, k, g.b2o) << cb) |
                     ((unsigned long long)j << (cb + 12));
        } else {
            key[r] = cm | ((unsigned long long)j << (cb + 12));
            if (j < nsamp) nflag |= 1u << r;
        }
    }
#pragma unroll
    for (int r = 0; r < RS_PER; r++)
        if (jw + (uint64_t)r * 64 + lane < nsamp)
            atomicAdd(&s_cnt[((nflag >> r) & 1u) ? 256u : ((uint32_t)key[r] & mask)], 1u);
    __syncthreads();
    /* thread d: digit d's offset over the earlier tiles; their sum is the
     * number of valid starts before this tile */
    const uint32_t pre = rs_lookback(status, tile, tid, s_cnt[tid], err);
    s_gofs[tid] = base[tid] + pre;
    uint32_t vbefore;
    (void)block_incl_scan_u32<RS_NT>(pre, s_w2, &vbefore);
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t *wc = s_wc + w * NC;
#pragma unroll
    for (int r = 0; r < RS_PER; r++) {
        const bool v = jw + (uint64_t)r * 64 + lane < nsamp;
        const bool nn = (nflag >> r) & 1u;
        const uint32_t c = nn ? 256u : ((uint32_t)key[r] & mask);
        const uint64_t bn = __ballot(nn);
        uint64_t peers = __ballot(v) & (nn ? bn : ~bn);
        for (uint32_t b = 0; b < bits; b++) {
            const uint64_t bb = __ballot((c >> b) & 1u);
            peers &= ((c >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t before = wc[c];
        rank[r] = before + (uint32_t)__popcll(peers & below);
        if (v && (peers & below) == 0) wc[c] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        /* thread d: the waves' offsets within class d and the class's start
         * in the tile; thread 0 also for the class of the starts with an N,
         * which starts after every valid key of the tile */
        const uint32_t d = tid;
        uint32_t tot = 0;
#pragma unroll
        for (int q = 0; q < RS_WAVES; q++) {
            const uint32_t c = s_wc[q * NC + d];
            s_wc[q * NC + d] = tot;
            tot += c;
        }
        if (d == 0) {
            uint32_t tn = 0;
#pragma unroll
            for (int q = 0; q < RS_WAVES; q++) {
                const uint32_t c = s_wc[q * NC + 256];
                s_wc[q * NC + 256] = tn;
                tn += c;
            }
        }
        uint32_t all;
        const uint32_t ds = block_incl_scan_u32<RS_NT>(tot, s_w, &all) - tot;
        s_dstart[d] = ds;
        if (d == 0) s_nval = all;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_PER; r++) {
        if (jw + (uint64_t)r * 64 + lane >= nsamp) continue;
        const bool nn = (nflag >> r) & 1u;
        const uint32_t c = nn ? 256u : ((uint32_t)key[r] & mask);
        s_k[(nn ? s_nval : s_dstart[c]) + s_wc[w * NC + c] + rank[r]] = key[r];
    }
    __syncthreads();
    const uint64_t tb = (uint64_t)tile * RS_TILE;
    const uint32_t nv = (uint32_t)(nsamp - tb < RS_TILE ? nsamp - tb : RS_TILE);
    const uint32_t nval = s_nval;
    const uint64_t nbase = *mp + (tb - vbefore);   /* starts with an N before this tile: tb - vbefore */
    for (uint32_t i = tid; i < nv; i += RS_NT) {
        const unsigned long long x = s_k[i];
        if (i < nval) {
            const uint32_t d = (uint32_t)x & mask;
            out[(uint64_t)s_gofs[d] + (i - s_dstart[d])] = x;
        } else {
            out[nbase + (i - nval)] = x;
        }
    }
}

/* Per chunk of `chunk` tiles and digit, the largest rl (the last code plus
 * one of the chunk's runs of that digit, 0 if none). */
__global__ void __launch_bounds__(256) k_rs_cmax(const uint32_t *__restrict__ rl, uint32_t ntiles, uint32_t chunk,
                                                 uint32_t *__restrict__ cmax)
{
    const uint32_t t0 = blockIdx.x * chunk, t1 = min(t0 + chunk, ntiles);
    uint32_t mx = 0;
    for (uint32_t t = t0; t < t1; t += 16) {
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = t + u < t1 ? rl[(uint64_t)(t + u) * 256 + threadIdx.x] : 0u;
#pragma unroll
        for (int u = 0; u < 16; u++) mx = max(mx, v[u]);
    }
    cmax[(uint64_t)blockIdx.x * 256 + threadIdx.x] = mx;
}

/* The bucket starts at the first key of each tile's digit run, after the
 * sort's last pass.  Its status rows hold every tile's inclusive count per
 * digit, so run (tile t, digit d) starts at g0 = base[d] + status[t - 1][d]
 * and is empty when that equals status[t][d].  The key before it in the
 * output is the last key of the previous non-empty run of digit d -- whose
 * code is the largest rl of digit d over the tiles before t (codes ascend
 * through the output) -- or, for the digit's first run, the last key of a
 * lower digit, whose code is read back through its position.  Codes
 * (previous, first] start at g0.  Workgroup = chunk of tiles, thread =
 * digit, tiles in order.  With no valid start every bucket starts at 0. */
__global__ void __launch_bounds__(256) k_rs_fix(const uint32_t *__restrict__ status, const uint32_t *__restrict__ base,
                                                const uint32_t *__restrict__ cmax, uint32_t ntiles, uint32_t chunk,
                                                const char *__restrict__ seq, const uint32_t *__restrict__ s2,
                                                const uint32_t *__restrict__ b3, uint64_t t0, SortFinish fin)
{
    const uint64_t m = *fin.m;
    const uint32_t d = threadIdx.x;
    if (m == 0) {
        for (uint64_t c = (uint64_t)blockIdx.x * 256 + d; c <= fin.nb; c += (uint64_t)gridDim.x * 256)
            fin.start[c] = 0;
        return;
    }
    uint32_t run = 0;                      /* largest rl of digit d before this tile */
    for (uint32_t c = 0; c < blockIdx.x; c += 16) {
        uint32_t v[16];
#pragma unroll
        for (int u = 0; u < 16; u++) v[u] = c + u < blockIdx.x ? cmax[(uint64_t)(c + u) * 256 + d] : 0u;
#pragma unroll
        for (int u = 0; u < 16; u++) run = max(run, v[u]);
    }
    const uint32_t ta = blockIdx.x * chunk, tz = min(ta + chunk, ntiles);
    uint32_t excl = ta ? status[(uint64_t)(ta - 1) * 256 + d] & RS_VAL : 0u;
    for (uint32_t tb = ta; tb < tz; tb += 16) {
        /* 16 tiles' status and run records loaded at once, then walked */
        uint32_t st[16], fr[16], lr[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const bool in = tb + u < tz;
            const uint64_t r = (uint64_t)(in ? tb + u : tb) * 256 + d;
            st[u] = status[r] & RS_VAL;
            fr[u] = fin.rf[r];
            lr[u] = fin.rl[r];
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const uint32_t incl = tb + u < tz ? st[u] : excl;   /* past the chunk: no run */
            if (incl != excl) {
                const uint64_t g0 = (uint64_t)base[d] + excl;
                if (g0 < m) {
                    const uint64_t code = fr[u];
                    uint64_t lo = ~0ull;   /* ~0: from code 0 */
                    if (excl) {
                        lo = run - 1u;
                    } else if (g0 > 0) {
                        uint32_t prev;
                        (void)start_code(seq, s2, b3, fin.positions[g0 - 1], t0, fin.k, &prev);
                        lo = prev;
                    }
                    for (uint64_t c = lo + 1; c <= code; c++) fin.start[c] = (uint32_t)g0;
                }
                run = max(run, lr[u]);
            }
            excl = incl;
        }
    }
}

inline uint32_t nblk(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

/* The sort's bookkeeping: digit counts and starts of up to four passes,
 * per-pass tile counters and look-back status, an error flag. */
struct RadixBufs {
    uint32_t *hist, *base, *ctr, *status, *err;
    uint64_t ntiles;
};

/* 8-bit passes over bits [lo, hi). */
static PassPlan plan_bits(uint32_t lo, uint32_t hi)
{
    PassPlan pl{};
    for (uint32_t sh = lo; sh < hi && pl.n < 4; sh += 8, pl.n++) {
        pl.sh[pl.n] = sh;
        pl.bits[pl.n] = hi - sh < 8 ? hi - sh : 8;
    }
    return pl;
}

/* The packed index's sort on the 2k code bits: 8-bit passes, and for
 * 2k <= 8 a pass 0 of no bits (k_pack_pass0 then only writes the keys)
 * before the one pass that orders them. */
static PassPlan plan_packed(uint32_t k)
{
    if (2 * k > 8) return plan_bits(0, 2 * k);
    PassPlan pl{};
    pl.n = 2;
    pl.bits[1] = 2 * k;
    return pl;
}

/* Sorts a[0..n) by the passes pl.sh/bits from pass p_first on, with b as the
 * other buffer; *res is the buffer that holds the result.  Nothing is
 * allocated.  counted: every pass's digit starts are in rb.base and the
 * look-back status and tile counters are zero (the kernels that wrote the
 * keys did that); otherwise they are cleared and counted here.  fin: the
 * last pass writes the index (positions, residuals, bucket starts) instead
 * of the sorted keys. */
static int radix_sort(unsigned long long *a, unsigned long long *b, uint64_t n, const PassPlan &pl, uint32_t p_first,
                      const RadixBufs &rb, int n_cu, hipStream_t st, unsigned long long **res, bool counted,
                      const SortFinish *fin)
{
    *res = a;
    if (n == 0 || pl.n == 0) return SWM_OK;
    if (pl.n > 4 || p_first > pl.n) return SWM_EINVAL;
    const uint64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
    if (ntiles > rb.ntiles || ntiles >= (1ull << 24)) return SWM_EINVAL;
    if (!counted) {
        SWM_HIP_CHECK(hipMemsetAsync(rb.ctr, 0, 4 * 4, st));
        SWM_HIP_CHECK(hipMemsetAsync(rb.status, 0, (size_t)pl.n * ntiles * 256 * 4, st));
        SWM_HIP_CHECK(hipMemsetAsync(rb.hist, 0, 4 * 256 * 4, st));
        k_rs_hist<<<(uint32_t)n_cu * 2, 256, 0, st>>>(a, n, pl, rb.hist);
        k_rs_bases<<<1, 256, 0, st>>>(rb.hist, pl.n, rb.base, nullptr);
        SWM_HIP_CHECK(hipGetLastError());
    }
    const SortFinish none{};
    unsigned long long *src = a, *dst = b;
    for (uint32_t pss = p_first; pss < pl.n; pss++) {
        const uint32_t sh = pl.sh[pss], bits = pl.bits[pss];
        uint32_t *stp = rb.status + (uint64_t)pss * ntiles * 256;
        if (fin && pss + 1 == pl.n)
            k_rs_pass<true><<<(uint32_t)ntiles, RS_NT, 0, st>>>(src, dst, n, sh, bits, rb.base + pss * 256, stp,
                                                                rb.ctr + pss, rb.err, *fin);
        else
            k_rs_pass<false><<<(uint32_t)ntiles, RS_NT, 0, st>>>(src, dst, n, sh, bits, rb.base + pss * 256, stp,
                                                                 rb.ctr + pss, rb.err, none);
        SWM_HIP_CHECK(hipGetLastError());
        unsigned long long *t = src; src = dst; dst = t;
    }
    *res = src;
    return SWM_OK;
}

/* Exclusive scan of n u32 with sums[ceil(n / SC_TILE)] as scratch. */
static int scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t *sums, hipStream_t st)
{
    if (n == 0) return SWM_OK;
    const uint32_t nt = (uint32_t)((n + SC_TILE - 1) / SC_TILE);
    k_sc_reduce<<<nt, SC_NT, 0, st>>>(in, n, sums);
    k_sc_down<<<nt, SC_NT, 0, st>>>(in, n, sums, out);
    SWM_HIP_CHECK(hipGetLastError());
    return SWM_OK;
}

struct DevBufs {
    void *p[32];
    int n = 0;
    ~DevBufs() { for (int i = 0; i < n; i++) (void)hipFree(p[i]); }
    template <typename T>
    bool get(T **ptr, size_t cnt)
    {
        if (hipMalloc((void **)ptr, (cnt ? cnt : 1) * sizeof(T)) != hipSuccess) return false;
        p[n++] = *ptr;
        return true;
    }
};

}  // namespace

extern "C" void swm_index_host_free(swm_index_host_t *ix)
{
    if (!ix) return;
    free(ix->base_2bits); free(ix->base_3bits); free(ix->hash_count);
    free(ix->hash_offset_end); free(ix->positions); free(ix->pos_lseeds);
    memset(ix, 0, sizeof(*ix));
}

extern "C" int swm_index_build_mem(const char *seq, uint64_t n, uint32_t k, swm_index_host_t *out, double *ms)
{
    if (!seq || !out || k < 2 || k > 14 || n < k || n >= 0xffffffffull) return SWM_EINVAL;
    memset(out, 0, sizeof(*out));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SWM_EDEVICE;
    const uint64_t nb = 1ull << (2 * k);
    const uint64_t nwords = (n + 31) / 32;
    const uint64_t nsamp = (n - k) / SWM_SEED_STEP + 1;
    const uint64_t t0 = (n > 64 + (uint64_t)k) ? ((n - 64 - k) & ~(uint64_t)3) : 0;
    uint32_t jbits = 0;
    while (jbits < 64 && (nsamp - 1) >> jbits) jbits++;
    /* SWM_INDEX_LARGE_PATH=1 takes the large-genome path at any size (its tests) */
    const char *lp = getenv("SWM_INDEX_LARGE_PATH");
    const bool packed = 2 * k + 12 + jbits <= 64 && !(lp && lp[0] == '1');
    hipStream_t st;
    SWM_HIP_CHECK(hipStreamCreate(&st));
    struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{st};
    DevBufs B;
    char *d_seq;
    uint32_t *b2, *b3, *s2, *cnt, *start, *end, *pos, *sums;
    unsigned long long *vin, *vout, *sorted = nullptr;
    uint64_t *d_m;
    uint16_t *ls;
    RadixBufs rb;
    rb.ntiles = (nsamp + RS_TILE - 1) / RS_TILE;
    const uint64_t nsums = (nb + 1 + SC_TILE - 1) / SC_TILE;
    int n_cu = 256;
    {
        int dev = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
    }
    /* every buffer up front (positions bounded by the sampled starts), so
     * the timed part has no allocation and no host round trip */
    if (!B.get(&d_seq, n + 64) || !B.get(&b2, 2 * nwords + 2 * SWM_WIN_PAIRS_PAD) ||
        !B.get(&b3, nwords + SWM_WIN_PAIRS_PAD) || !B.get(&s2, 2 * nwords + 2 * SWM_WIN_PAIRS_PAD) ||
        !B.get(&cnt, nb + 1) || !B.get(&start, nb + 1) || !B.get(&end, nb + 1) || !B.get(&vin, nsamp) ||
        !B.get(&vout, nsamp) || !B.get(&d_m, 1) || !B.get(&pos, nsamp + 64) || !B.get(&ls, nsamp + 64) ||
        !B.get(&sums, nsums) || !B.get(&rb.hist, 4 * 256 + 8) || !B.get(&rb.base, 4 * 256) ||
        !B.get(&rb.status, 4 * rb.ntiles * 256) || !B.get(&rb.err, 1))
        return SWM_ENOMEM;
    rb.ctr = rb.hist + 4 * 256;             /* cleared with the counts by k_encode */
    SWM_HIP_CHECK(hipMemsetAsync(rb.err, 0, 4, st));
    /* the sort's last-pass run records: chunks of tiles, at most 256 chunks */
    const uint32_t chunk = (uint32_t)(rb.ntiles <= 64 * 256 ? 64 : (rb.ntiles + 255) / 256);
    const uint32_t nchunk = (uint32_t)((rb.ntiles + chunk - 1) / chunk);
    uint32_t *rf = nullptr, *rl = nullptr, *cmax = nullptr;
    if (packed && (!B.get(&rf, rb.ntiles * 256) || !B.get(&rl, rb.ntiles * 256) ||
                   !B.get(&cmax, (uint64_t)nchunk * 256)))
        return SWM_ENOMEM;
    SWM_HIP_CHECK(hipMemsetAsync(d_seq, 0, n + 64, st));
    SWM_HIP_CHECK(hipMemcpyAsync(d_seq, seq, n, hipMemcpyHostToDevice, st));
    SWM_HIP_CHECK(hipMemsetAsync(b2, 0, (2 * nwords + 2 * SWM_WIN_PAIRS_PAD) * 4, st));
    SWM_HIP_CHECK(hipMemsetAsync(b3, 0, (nwords + SWM_WIN_PAIRS_PAD) * 4, st));
    SWM_HIP_CHECK(hipMemsetAsync(s2, 0, (2 * nwords + 2 * SWM_WIN_PAIRS_PAD) * 4, st));
    SWM_HIP_CHECK(hipMemsetAsync(cnt, 0, (nb + 1) * 4, st));
    hipEvent_t e0, e1;
    SWM_HIP_CHECK(hipEventCreate(&e0));
    SWM_HIP_CHECK(hipEventCreate(&e1));
    struct EG { hipEvent_t a, b; ~EG() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); } } eg{e0, e1};
    SWM_HIP_CHECK(hipEventRecord(e0, st));
    k_encode<<<nblk(nwords, 256), 256, 0, st>>>(d_seq, n, nwords, b2, b3, s2, rb.hist, 4 * 256 + 8);
    SWM_HIP_CHECK(hipGetLastError());
    if (packed) {
        /* every pass's digit counts from the genome words, the keys written
         * by the first pass of their sort on the code bits, the other
         * passes, the last of which writes positions, residuals and starts */
        const PassPlan pl = plan_packed(k);
        const uint32_t hgrid = (uint32_t)std::min<uint64_t>((uint64_t)n_cu * 4, rb.ntiles);
        k_pack_hist<<<hgrid, 256, 0, st>>>(d_seq, nsamp, k, t0, s2, b3, b2, pl, rb.hist, rb.status,
                                           (uint64_t)pl.n * rb.ntiles * 256);
        k_rs_bases<<<1, 256, 0, st>>>(rb.hist, pl.n, rb.base, d_m);
        k_pack_pass0<<<(uint32_t)rb.ntiles, RS_NT, 0, st>>>(d_seq, nsamp, k, t0, s2, b3, b2, pl.bits[0], rb.base,
                                                            d_m, rb.status, rb.ctr, rb.err, vin);
        SWM_HIP_CHECK(hipGetLastError());
        const SortFinish fin{d_m, k, nb, start, pos, ls, rf, rl};
        int rc = radix_sort(vin, vout, nsamp, pl, 1, rb, n_cu, st, &sorted, true, &fin);
        if (rc) return rc;
        k_rs_cmax<<<nchunk, 256, 0, st>>>(rl, (uint32_t)rb.ntiles, chunk, cmax);
        k_rs_fix<<<nchunk, 256, 0, st>>>(rb.status + (uint64_t)(pl.n - 1) * rb.ntiles * 256,
                                         rb.base + (pl.n - 1) * 256, cmax, (uint32_t)rb.ntiles, chunk, d_seq, s2, b3,
                                         t0, fin);
        k_counts<<<nblk(nb + 1, 256), 256, 0, st>>>(nb, start, cnt, end);
        SWM_HIP_CHECK(hipGetLastError());
    } else {
        /* stable sort by code (the counting sort's order, positions
         * ascending), counts from the sorted runs, offsets by scan */
        k_keys<<<nblk(nsamp, 256), 256, 0, st>>>(d_seq, nsamp, k, vin);
        SWM_HIP_CHECK(hipGetLastError());
        int rc = radix_sort(vin, vout, nsamp, plan_bits(32, 32 + 2 * k + 1), 0, rb, n_cu, st, &sorted, false,
                            nullptr);
        if (rc) return rc;
        k_runs<<<nblk(nsamp, 256), 256, 0, st>>>(nsamp, (uint32_t)nb, sorted, cnt);
        SWM_HIP_CHECK(hipGetLastError());
        if ((rc = scan_u32(cnt, start, nb + 1, sums, st))) return rc;
        k_m_from_start<<<1, 1, 0, st>>>(start, nb, d_m);
        k_split<<<nblk(nsamp, 256), 256, 0, st>>>(d_m, sorted, b2, k, pos, ls);
        k_end_offsets<<<nblk(nb + 1, 256), 256, 0, st>>>(nb, start, end);
    }
    k_tail<<<1, 64, 0, st>>>(d_seq, n, k, t0, b2, b3, s2, cnt, start, d_m, pos, ls, end);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(e1, st));
    uint64_t M = 0;
    uint32_t sort_err = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(&M, d_m, 8, hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipMemcpyAsync(&sort_err, rb.err, 4, hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    if (M > nsamp || sort_err) return SWM_EDEVICE;

    out->k = k;
    out->n_bases = n;
    out->n_words = nwords;
    out->n_pos = M;
    out->base_2bits = (uint32_t *)malloc(2 * nwords * 4);
    out->base_3bits = (uint32_t *)malloc(nwords * 4);
    out->hash_count = (uint32_t *)malloc(nb * 4);
    out->hash_offset_end = (uint32_t *)malloc((nb + 1) * 4);
    out->positions = (uint32_t *)malloc((M + 1) * 4);
    out->pos_lseeds = (uint16_t *)malloc((M + 1) * 2);
    if (!out->base_2bits || !out->base_3bits || !out->hash_count || !out->hash_offset_end || !out->positions ||
        !out->pos_lseeds) {
        swm_index_host_free(out);
        return SWM_ENOMEM;
    }
    if (hipMemcpyAsync(out->base_2bits, b2, 2 * nwords * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(out->base_3bits, b3, nwords * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(out->hash_count, cnt, nb * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(out->hash_offset_end, end, (nb + 1) * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (M && hipMemcpyAsync(out->positions, pos, M * 4, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        (M && hipMemcpyAsync(out->pos_lseeds, ls, M * 2, hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
        swm_index_host_free(out);
        return SWM_EDEVICE;
    }
    if (ms) {
        float f = 0;
        (void)hipEventElapsedTime(&f, e0, e1);
        *ms = f;
    }
    return SWM_OK;
}
