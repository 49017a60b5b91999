/*
 * swm_index.hip -- 12-mer hash index build on the GPU, byte-identical to
 * serial_index.c under HOST_BUILD (construct_hash_index, :359-445).
 *
 *   k_encode   one thread per 32-base word: bit planes (.base_2bits), N mask
 *              (.base_3bits) and the build-only 2-bit seed words
 *              (get_seed/get_base, serial_index.c:146-235), including the
 *              final flush that zeroes the last word when the genome length
 *              is a multiple of the word size (:423-426)
 *   k_pack_*   the sampled starts p = 0, 4, 8, ... <= N-k whose k-mer has no
 *              N (the histogram's starts, :253-269), compacted in genome
 *              order (count, scan, write), each as one 64-bit key
 *              code | residual << 2k | (p/4) << (2k+12): the k-mer code from
 *              the 2-bit seed words, the N test from the mask, the 12-bit
 *              residual (get_seed_from_enc_bits :280-331) from the planes;
 *              the last starts (the tail below) from the sequence itself
 *   sort       stable LSD radix sort on the code bits only (3 passes for
 *              k=12): the counting sort's bucket order with positions
 *              ascending (:343-352)
 *   k_finish64 .positions and .pos_lseeds in bucket order, and the bucket
 *              starts (store_positions' offsets :337-340) from the first
 *              key of each code
 *   k_counts   .hash_count and the END offsets from the starts
 *   k_tail     one wave replays the scatter for the last starts, whose
 *              scatter reads the zeroed words, so that slots, overflow and
 *              the END offsets of .hash_offset (Q19) match the serial build
 * Genomes whose key would not fit 64 bits (more than 2^(52-2k) sampled
 * starts) take the pair sort: key = code, value = residual << 32 | p.
 */
#include "swm_internal.h"

#include <hipcub/hipcub.hpp>
#include <stdlib.h>
#include <string.h>

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

__global__ void __launch_bounds__(256) k_encode(const char *__restrict__ seq, uint64_t n, uint64_t nwords,
                                                uint32_t *__restrict__ b2, uint32_t *__restrict__ b3,
                                                uint32_t *__restrict__ s2)
{
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
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
__device__ __forceinline__ uint32_t lseed_at(const uint32_t *__restrict__ b2, uint64_t pos, uint32_t k)
{
    uint64_t s_pos = pos >= 3 ? pos - 3 : 0;
    uint64_t wi = s_pos >> 5;
    uint32_t sh = (uint32_t)s_pos & 31u;
    uint32_t fl = __builtin_amdgcn_alignbit(b2[2 * wi + 2], b2[2 * wi], sh) & 7u;
    uint32_t fh = __builtin_amdgcn_alignbit(b2[2 * wi + 3], b2[2 * wi + 1], sh) & 7u;
    s_pos = pos + k;
    wi = s_pos >> 5;
    sh = (uint32_t)s_pos & 31u;
    uint32_t bl = __builtin_amdgcn_alignbit(b2[2 * wi + 2], b2[2 * wi], sh) & 7u;
    uint32_t bh = __builtin_amdgcn_alignbit(b2[2 * wi + 3], b2[2 * wi + 1], sh) & 7u;
    fl |= bl << 3;
    fh |= bh << 3;
    return fl | (fh << 6);
}

/* One thread per sampled start p = 0, 4, 8, ... <= N-k (in genome order, so
 * the plane reads are coalesced): the k-mer code and N-freeness of the
 * histogram (:253-269) as the sort key (4^k for a start with an N), and the
 * value (residual << 32 | p) that the stable sort carries into bucket
 * order. */
__global__ void __launch_bounds__(256) k_keys(const char *__restrict__ seq, uint64_t nsamp, uint32_t k,
                                              const uint32_t *__restrict__ b2, uint32_t *__restrict__ keys,
                                              unsigned long long *__restrict__ vals)
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
    keys[j] = valid ? code : (1u << (2 * k));
    vals[j] = ((unsigned long long)lseed_at(b2, p, k) << 32) | (uint32_t)p;
}

#define PACK_NT 256
#define PACK_PER 4
#define PACK_TILE (PACK_NT * PACK_PER)

/* The histogram's code and N test of start p: from the seed words and the
 * mask for p < t0, from the sequence for the last starts (whose words the
 * final flush may have zeroed). */
__device__ __forceinline__ bool start_code(const char *__restrict__ seq, const uint32_t *__restrict__ s2,
                                           const uint32_t *__restrict__ b3, uint64_t p, uint64_t t0, uint32_t k,
                                           uint32_t *code)
{
    if (p < t0) {
        const uint64_t sw = p >> 4, mw = p >> 5;
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

/* Pass 1: valid starts per tile of PACK_TILE starts. */
__global__ void __launch_bounds__(PACK_NT) k_pack_count(const char *__restrict__ seq, uint64_t nsamp, uint32_t k,
                                                        uint64_t t0, const uint32_t *__restrict__ s2,
                                                        const uint32_t *__restrict__ b3, uint32_t *__restrict__ tcnt)
{
    __shared__ uint32_t s_w[PACK_NT / 64];
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < PACK_PER; q++) {
        const uint64_t j = (uint64_t)blockIdx.x * PACK_TILE + (uint64_t)q * PACK_NT + threadIdx.x;
        uint32_t code;
        if (j < nsamp && start_code(seq, s2, b3, j * SWM_SEED_STEP, t0, k, &code)) c++;
    }
    uint32_t tot;
    (void)block_incl_scan_u32<PACK_NT>(c, s_w, &tot);
    if (threadIdx.x == 0) tcnt[blockIdx.x] = tot;
}

/* Pass 2: the keys of the valid starts at their rank (genome order).  A
 * tile is PACK_PER rows of PACK_NT consecutive starts, so each row's stores
 * are contiguous; one scan of four 16-bit row counters ranks all rows. */
__global__ void __launch_bounds__(PACK_NT) k_pack_write(const char *__restrict__ seq, uint64_t nsamp, uint32_t k,
                                                        uint64_t t0, const uint32_t *__restrict__ s2,
                                                        const uint32_t *__restrict__ b3,
                                                        const uint32_t *__restrict__ b2,
                                                        const uint32_t *__restrict__ toff,
                                                        const uint32_t *__restrict__ tcnt, uint32_t ntiles,
                                                        unsigned long long *__restrict__ keys)
{
    __shared__ uint64_t s_w[PACK_NT / 64];
    const uint64_t tile0 = (uint64_t)blockIdx.x * PACK_TILE;
    uint32_t code[PACK_PER];
    uint64_t flags = 0;
#pragma unroll
    for (int q = 0; q < PACK_PER; q++) {
        const uint64_t j = tile0 + (uint64_t)q * PACK_NT + threadIdx.x;
        const bool ok = j < nsamp && start_code(seq, s2, b3, j * SWM_SEED_STEP, t0, k, &code[q]);
        flags |= (uint64_t)(ok ? 1u : 0u) << (16 * q);
    }
    /* inclusive scan of the packed row counters over the workgroup */
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    uint64_t v = flags;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d, 64);
        const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d, 64);
        if (lane >= (uint32_t)d) v += ((uint64_t)hi << 32) | lo;
    }
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    uint64_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < PACK_NT / 64; i++) {
        const uint64_t x = s_w[i];
        if ((uint32_t)i < w) off += x;
        tot += x;
    }
    const uint64_t excl = v + off - flags;
    /* starts with an N go after all valid ones, with the largest code, so
     * the sort leaves them at the end (M = valid starts) */
    const uint64_t M = (uint64_t)toff[ntiles - 1] + tcnt[ntiles - 1];
    const uint64_t tbase = toff[blockIdx.x];
    const unsigned long long cm = (1ull << (2 * k)) - 1ull;
    uint64_t rows_before = 0;
#pragma unroll
    for (int q = 0; q < PACK_PER; q++) {
        const uint64_t j = tile0 + (uint64_t)q * PACK_NT + threadIdx.x;
        const uint64_t rank = rows_before + ((excl >> (16 * q)) & 0xffffu);   /* valid before j in the tile */
        rows_before += (tot >> (16 * q)) & 0xffffu;
        if (j >= nsamp) continue;
        if ((flags >> (16 * q)) & 1u) {
            const uint64_t res = lseed_at(b2, j * SWM_SEED_STEP, k);
            keys[tbase + rank] = (unsigned long long)code[q] | (res << (2 * k)) | (j << (2 * k + 12));
        } else {
            keys[M + (tile0 - tbase) + ((j - tile0) - rank)] = cm | (j << (2 * k + 12));
        }
    }
}

/* From the sorted 64-bit keys (code in the low 2k bits): .positions and
 * .pos_lseeds in bucket order, and the bucket starts -- the key at the
 * first index of each code sets the start of that code and of every empty
 * code below it down to the previous key's, the end closes the codes above
 * the last key with M.  Every start is written exactly once (start[nb] = M),
 * so no count pass and no scan are needed. */
__global__ void __launch_bounds__(256) k_finish64(const uint64_t *__restrict__ mp, uint32_t k, uint64_t nb,
                                                  const unsigned long long *__restrict__ key,
                                                  uint32_t *__restrict__ start, uint32_t *__restrict__ positions,
                                                  uint16_t *__restrict__ lseeds)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = *mp;
    if (i > m) return;
    const unsigned long long cm = (1ull << (2 * k)) - 1ull;
    uint64_t hi = nb;
    if (i < m) {
        const unsigned long long x = key[i];
        positions[i] = (uint32_t)((x >> (2 * k + 12)) * SWM_SEED_STEP);
        lseeds[i] = (uint16_t)((x >> (2 * k)) & 0xfffu);
        hi = x & cm;
    }
    const uint64_t lo = i > 0 ? (key[i - 1] & cm) : ~0ull;   /* ~0: from code 0 */
    if (hi == lo) return;
    for (uint64_t c = lo + 1; c <= hi; c++) start[c] = (uint32_t)i;      /* codes (lo, hi] start at i */
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

/* The number of valid starts (M) from the last tile offset and count. */
__global__ void k_total(const uint32_t *__restrict__ toff, const uint32_t *__restrict__ tcnt, uint32_t ntiles,
                        uint64_t *__restrict__ m)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *m = ntiles ? (uint64_t)toff[ntiles - 1] + tcnt[ntiles - 1] : 0;
}

/* Bucket counts from the sorted keys: the thread at the end of each run of
 * one valid code gallops back to the run's start (runs average a few
 * entries; long ones cost log steps). */
__global__ void __launch_bounds__(256) k_runs(uint64_t nsamp, uint32_t nb, const uint32_t *__restrict__ key,
                                              uint32_t *__restrict__ count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nsamp) return;
    const uint32_t c = key[i];
    if (c >= nb || (i + 1 < nsamp && key[i + 1] == c)) return;
    uint64_t lo = i, step = 1;             /* key[lo] == c */
    while (step <= lo && key[lo - step] == c) { lo -= step; step <<= 1; }
    uint64_t a = step <= lo ? lo - step : 0;   /* key[a] != c or a == 0 */
    if (key[a] == c) lo = a;
    else {
        while (lo - a > 1) {               /* key[a] != c, key[lo] == c */
            const uint64_t mid = a + (lo - a) / 2;
            if (key[mid] == c) lo = mid; else a = mid;
        }
    }
    count[c] = (uint32_t)(i + 1 - lo);
}

/* positions and residuals in bucket order */
__global__ void __launch_bounds__(256) k_split(const uint64_t *__restrict__ mp, const unsigned long long *__restrict__ v,
                                               uint32_t *__restrict__ positions, uint16_t *__restrict__ lseeds)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *mp) return;
    const unsigned long long x = v[i];
    positions[i] = (uint32_t)x;
    lseeds[i] = (uint16_t)(x >> 32);
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

inline uint32_t nblk(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

struct DevBufs {
    void *p[16];
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
    const bool packed = 2 * k + 12 + jbits <= 64;
    const uint32_t ntiles = (uint32_t)((nsamp + PACK_TILE - 1) / PACK_TILE);
    hipStream_t st;
    SWM_HIP_CHECK(hipStreamCreate(&st));
    struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{st};
    DevBufs B;
    char *d_seq;
    uint32_t *b2, *b3, *s2, *cnt, *start, *end, *pos, *kin = nullptr, *kout = nullptr, *tcnt = nullptr, *toff = nullptr;
    unsigned long long *vin, *vout;
    uint64_t *d_m;
    uint16_t *ls;
    /* every buffer up front (positions bounded by the sampled starts), so
     * the timed part has no allocation and no host round trip */
    if (!B.get(&d_seq, n + 64) || !B.get(&b2, 2 * nwords + 2 * SWM_WIN_PAIRS_PAD) ||
        !B.get(&b3, nwords + SWM_WIN_PAIRS_PAD) || !B.get(&s2, 2 * nwords + 2 * SWM_WIN_PAIRS_PAD) ||
        !B.get(&cnt, nb + 1) || !B.get(&start, nb + 1) || !B.get(&end, nb + 1) || !B.get(&vin, nsamp) ||
        !B.get(&vout, nsamp) || !B.get(&d_m, 1) || !B.get(&pos, nsamp + 64) || !B.get(&ls, nsamp + 64))
        return SWM_ENOMEM;
    if (packed ? (!B.get(&tcnt, ntiles) || !B.get(&toff, ntiles)) : (!B.get(&kin, nsamp) || !B.get(&kout, nsamp)))
        return SWM_ENOMEM;
    size_t t1 = 0, t2 = 0, t3 = 0;
    SWM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, cnt, start, nb + 1, st));
    if (packed) {
        SWM_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, t2, vin, vout, nsamp, 0, (int)(2 * k), st));
        SWM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t3, tcnt, toff, ntiles, st));
    } else {
        SWM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, kin, kout, vin, vout, nsamp, 0,
                                                         (int)(2 * k + 1), st));
    }
    size_t tb = t1 > t2 ? t1 : t2;
    if (t3 > tb) tb = t3;
    void *tmp;
    if (!B.get((char **)&tmp, tb)) return SWM_ENOMEM;
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
    k_encode<<<nblk(nwords, 256), 256, 0, st>>>(d_seq, n, nwords, b2, b3, s2);
    SWM_HIP_CHECK(hipGetLastError());
    if (packed) {
        /* valid starts compacted in genome order as 64-bit keys, a sort on
         * the code bits, counts, offsets, positions and residuals */
        k_pack_count<<<ntiles, PACK_NT, 0, st>>>(d_seq, nsamp, k, t0, s2, b3, tcnt);
        SWM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, t3, tcnt, toff, ntiles, st));
        k_pack_write<<<ntiles, PACK_NT, 0, st>>>(d_seq, nsamp, k, t0, s2, b3, b2, toff, tcnt, ntiles, vin);
        k_total<<<1, 1, 0, st>>>(toff, tcnt, ntiles, d_m);
        SWM_HIP_CHECK(hipGetLastError());
        SWM_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(tmp, t2, vin, vout, nsamp, 0, (int)(2 * k), st));
        k_finish64<<<nblk(nsamp + 1, 256), 256, 0, st>>>(d_m, k, nb, vout, start, pos, ls);
        k_counts<<<nblk(nb + 1, 256), 256, 0, st>>>(nb, start, cnt, end);
        SWM_HIP_CHECK(hipGetLastError());
    } else {
        /* stable sort by code (the counting sort's order, positions
         * ascending), counts from the sorted runs, offsets by scan */
        k_keys<<<nblk(nsamp, 256), 256, 0, st>>>(d_seq, nsamp, k, b2, kin, vin);
        SWM_HIP_CHECK(hipGetLastError());
        SWM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, t2, kin, kout, vin, vout, nsamp, 0, (int)(2 * k + 1),
                                                         st));
        k_runs<<<nblk(nsamp, 256), 256, 0, st>>>(nsamp, (uint32_t)nb, kout, cnt);
        SWM_HIP_CHECK(hipGetLastError());
        SWM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, t1, cnt, start, nb + 1, st));
        k_m_from_start<<<1, 1, 0, st>>>(start, nb, d_m);
        k_split<<<nblk(nsamp, 256), 256, 0, st>>>(d_m, vout, pos, ls);
        k_end_offsets<<<nblk(nb + 1, 256), 256, 0, st>>>(nb, start, end);
    }
    k_tail<<<1, 64, 0, st>>>(d_seq, n, k, t0, b2, b3, s2, cnt, start, d_m, pos, ls, end);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(e1, st));
    uint64_t M = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(&M, d_m, 8, hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    if (M > nsamp) return SWM_EDEVICE;

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
