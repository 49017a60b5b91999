/*
 * swm_index.hip -- 12-mer hash index build on the GPU, byte-identical to
 * serial_index.c under HOST_BUILD (construct_hash_index, :359-445).
 *
 *   k_encode   one thread per 32-base word: bit planes (.base_2bits), N mask
 *              (.base_3bits) and the build-only 2-bit seed words
 *              (get_seed/get_base, serial_index.c:146-235), including the
 *              final flush that zeroes the last word when the genome length
 *              is a multiple of the word size (:423-426)
 *   k_keys     one thread per sampled start p = 0, 4, 8, ... <= N-k, in
 *              genome order: the k-mer code and N-freeness of the histogram
 *              (:253-269) as the key, the 12-bit residual code
 *              (get_seed_from_enc_bits :280-331) with p as the value
 *   sort       stable LSD radix sort of (code, residual|p): the counting
 *              sort's bucket order with positions ascending (:343-352)
 *   k_runs     bucket counts (.hash_count) from the sorted runs, no atomics
 *   scan       exclusive scan of the counts (store_positions :337-340)
 *   k_split    .positions and .pos_lseeds in bucket order
 *   k_tail     one thread replays the scatter for the last starts, whose
 *              scatter reads the zeroed words, so that slots, overflow and
 *              the END offsets of .hash_offset (Q19) match the serial build
 */
#include "swm_internal.h"

#include <hipcub/hipcub.hpp>
#include <stdlib.h>
#include <string.h>

namespace {

__device__ __forceinline__ int base_code(char raw)
{
    char c = (raw >= 'a' && raw <= 'z') ? (char)(raw - 32) : raw;   /* toupper, serial_index.c:152 */
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
    }
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
__global__ void __launch_bounds__(256) k_split(uint64_t M, const unsigned long long *__restrict__ v,
                                               uint32_t *__restrict__ positions, uint16_t *__restrict__ lseeds)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const unsigned long long x = v[i];
    positions[i] = (uint32_t)x;
    lseeds[i] = (uint16_t)(x >> 32);
}

__global__ void k_end_offsets(uint64_t nb, const uint32_t *__restrict__ start, uint32_t *__restrict__ end)
{
    uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nb) end[c] = start[c + 1];
    if (c == nb) end[nb] = start[nb];
}

#define TAIL_MAX 64

/* Serial replay of store_positions (serial_index.c:343-352) for the starts
 * p >= t0, whose scatter code and N check read the (possibly zeroed) last
 * words.  All earlier starts read the same bits as the histogram. */
__global__ void k_tail(const char *__restrict__ seq, uint64_t n, uint32_t k, uint64_t t0,
                       const uint32_t *__restrict__ b2, const uint32_t *__restrict__ b3,
                       const uint32_t *__restrict__ s2, const uint32_t *__restrict__ count,
                       const uint32_t *__restrict__ start, uint64_t M, uint32_t *__restrict__ positions,
                       uint16_t *__restrict__ lseeds, uint32_t *__restrict__ end)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    uint32_t tp[TAIL_MAX], tcode[TAIL_MAX], scode[TAIL_MAX];
    bool tvalid[TAIL_MAX], svalid[TAIL_MAX];
    uint32_t nt = 0;
    const uint32_t smask = (uint32_t)((1ull << (2 * k)) - 1), fmask = (uint32_t)((1ull << k) - 1);
    for (uint64_t p = t0; p + k <= n && nt < TAIL_MAX; p += SWM_SEED_STEP) {
        uint32_t code = 0;
        bool v = true;
        for (uint32_t i = 0; i < k; i++) {
            int c = base_code(seq[p + i]);
            if (c < 0) v = false;
            code |= (uint32_t)(c < 0 ? 0 : c) << (2 * i);
        }
        tp[nt] = (uint32_t)p;
        tcode[nt] = code;
        tvalid[nt] = v;
        uint32_t sw = __builtin_amdgcn_alignbit(s2[p / 16 + 1], s2[p / 16], (uint32_t)((p % 16) * 2)) & smask;
        uint32_t fl = __builtin_amdgcn_alignbit(b3[p / 32 + 1], b3[p / 32], (uint32_t)(p % 32)) & fmask;
        scode[nt] = sw;
        svalid[nt] = (fl == 0);
        nt++;
    }
    /* main_cnt(c) = count[c] - (tail starts whose true code is c) */
    auto tail_true = [&](uint32_t c) {
        uint32_t t = 0;
        for (uint32_t i = 0; i < nt; i++) t += (tvalid[i] && tcode[i] == c) ? 1u : 0u;
        return t;
    };
    /* 1) clear the slots the bucket sort gave to tail starts (the serial
     *    scatter writes them only through the replay below) */
    for (uint32_t i = 0; i < nt; i++) {
        if (!tvalid[i]) continue;
        uint32_t c = tcode[i];
        uint32_t rank = 0;
        for (uint32_t q = 0; q < i; q++) rank += (tvalid[q] && tcode[q] == c) ? 1u : 0u;
        uint64_t slot = (uint64_t)start[c] + (count[c] - tail_true(c)) + rank;
        if (slot < M) { positions[slot] = 0; lseeds[slot] = 0; }
    }
    /* 2) replay the serial scatter for the tail, in start order */
    for (uint32_t i = 0; i < nt; i++) {
        if (!svalid[i]) continue;
        uint32_t c = scode[i];
        uint32_t prior = 0;
        for (uint32_t q = 0; q < i; q++) prior += (svalid[q] && scode[q] == c) ? 1u : 0u;
        uint64_t slot = (uint64_t)start[c] + (count[c] - tail_true(c)) + prior;
        if (slot < M) {
            positions[slot] = tp[i];
            lseeds[slot] = (uint16_t)lseed_at(b2, tp[i], k);
        }
    }
    /* 3) END offsets of every bucket the tail touches */
    for (uint32_t i = 0; i < 2 * nt; i++) {
        bool isv = (i < nt) ? tvalid[i] : svalid[i - nt];
        if (!isv) continue;
        uint32_t c = (i < nt) ? tcode[i] : scode[i - nt];
        uint32_t writes = 0;
        for (uint32_t q = 0; q < nt; q++) writes += (svalid[q] && scode[q] == c) ? 1u : 0u;
        end[c] = start[c] + (count[c] - tail_true(c)) + writes;
    }
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
    hipStream_t st;
    SWM_HIP_CHECK(hipStreamCreate(&st));
    struct SG { hipStream_t s; ~SG() { (void)hipStreamDestroy(s); } } sg{st};
    DevBufs B;
    char *d_seq;
    uint32_t *b2, *b3, *s2, *cnt, *start, *end, *kin, *kout, *pos;
    unsigned long long *vin, *vout;
    uint16_t *ls;
    if (!B.get(&d_seq, n + 64) || !B.get(&b2, 2 * nwords + 2 * SWM_WIN_PAIRS_PAD) ||
        !B.get(&b3, nwords + SWM_WIN_PAIRS_PAD) || !B.get(&s2, 2 * nwords + 2 * SWM_WIN_PAIRS_PAD) ||
        !B.get(&cnt, nb + 1) || !B.get(&start, nb + 1) || !B.get(&end, nb + 1) || !B.get(&kin, nsamp) ||
        !B.get(&kout, nsamp) || !B.get(&vin, nsamp) || !B.get(&vout, nsamp))
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
    k_encode<<<nblk(nwords, 256), 256, 0, st>>>(d_seq, n, nwords, b2, b3, s2);
    k_keys<<<nblk(nsamp, 256), 256, 0, st>>>(d_seq, nsamp, k, b2, kin, vin);
    SWM_HIP_CHECK(hipGetLastError());
    size_t t1 = 0, t2 = 0;
    SWM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, cnt, start, nb + 1, st));
    SWM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, kin, kout, vin, vout, nsamp, 0, (int)(2 * k + 1), st));
    void *tmp;
    if (!B.get((char **)&tmp, t1 > t2 ? t1 : t2)) return SWM_ENOMEM;
    /* stable sort by code (the counting sort's order, positions ascending),
     * counts from the sorted runs, offsets by scan */
    SWM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, t2, kin, kout, vin, vout, nsamp, 0, (int)(2 * k + 1), st));
    k_runs<<<nblk(nsamp, 256), 256, 0, st>>>(nsamp, (uint32_t)nb, kout, cnt);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, t1, cnt, start, nb + 1, st));
    uint32_t M32 = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(&M32, start + nb, 4, hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    const uint64_t M = M32;
    if (!B.get(&pos, M + 64) || !B.get(&ls, M + 64)) return SWM_ENOMEM;
    k_split<<<nblk(M ? M : 1, 256), 256, 0, st>>>(M, vout, pos, ls);
    k_end_offsets<<<nblk(nb + 1, 256), 256, 0, st>>>(nb, start, end);
    const uint64_t t0 = (n > 64 + (uint64_t)k) ? ((n - 64 - k) & ~(uint64_t)3) : 0;
    k_tail<<<1, 1, 0, st>>>(d_seq, n, k, t0, b2, b3, s2, cnt, start, M, pos, ls, end);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipEventRecord(e1, st));

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
