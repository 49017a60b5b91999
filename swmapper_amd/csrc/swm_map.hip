/*
 * swm_map.hip -- MI355X mapping hot path: seed lookup, candidate batches,
 * 15-mer filter, banded Myers verification and record emission.
 *
 * The path replaces find_positions (mapper_cpe.c:2592-2663) behind
 * swm_map_batch*.  One launch per stage on the context's stream:
 *
 *   k_seed_fixed<12> (k_seed for other k)  64 reads per workgroup
 *              staged in LDS, one thread per (read, strand): read encoding
 *              (applied from a per-length
 *              recipe), 12-mer hash-offset gathers, long-seed ranking,
 *              residuals (mapper_cpe.c:1246-1401, 2451-2499)
 *   k_plan_pair  one wave per read, both strands' light items: the one batch
 *              of alignment_rmdup (mapper_cpe.c:2124-2426) -- the selected
 *              12-mers' positions fetched 64 lanes at a time, filtered by the
 *              15-mer test (:2090-2123), shifted to read starts
 *              (:1893-1917), sorted and uniqued (block_heap_sort's result,
 *              :1999-2089); heavy items (a 12-mer with more than one chunk)
 *              go to k_plan_heavy (seed-heap replay) + k_plan_fetch;
 *              k_plan_plain is the whole state machine one thread per item
 *   k_bigsort  one wave per large batch: register bitonic sort + unique
 *   k_compact  256 (read, strand) items per workgroup, copied flat: dense
 *              candidate list (k_compact_heavy: one heavy item per workgroup)
 *   k_myers2   two candidates per lane (e <= 7) / k_myers one: window
 *              fetch, banded Myers (semantics mapper_cpe.c:519-660 == :668-976),
 *              then for each passing candidate the Hamming check of
 *              align_reads (:252-287) on the same window words
 *   k_trace    one lane per passing candidate that failed the Hamming check:
 *              the traceback of align_reads (:288-509)
 *   k_emit     one thread per read (k_emit_heavy one wave per heavy read):
 *              duplicate rule and flush state machine (:475-477, 1790, 1858,
 *              1437-1507), per-read reduction
 */
#include "swm_internal.h"

#include <algorithm>
#include <new>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <vector>

#include "swm_recipe.h"

namespace {

/* ---------------------------------------------------------------- helpers */

__device__ __forceinline__ uint32_t shl32(uint32_t x, int s) { return (s < 0 || s >= 32) ? 0u : (x << s); }

/* Plane word (0 = low bit plane, 1 = high) of the genome starting at absolute
 * base g.  The device copy of .base_2bits carries SWM_WIN_PAIRS_PAD zero
 * pairs after the file, so windows past the end read 0; bases before 0 read 0. */
__device__ __forceinline__ uint32_t genome_word(const uint32_t *__restrict__ b2, int64_t g, int plane)
{
    /* without branches, so that a caller's loads issue together: a start
     * before the genome reads word 0 (the word a negative start shifts) */
    const uint64_t gc = g < 0 ? 0u : (uint64_t)g;
    const uint64_t wi = gc >> 5;
    const uint32_t a = b2[2 * wi + plane], b = b2[2 * wi + 2 + plane];
    const uint32_t in = __builtin_amdgcn_alignbit(b, a, (uint32_t)gc & 31u);
    return g >= 0 ? in : (g <= -32 ? 0u : shl32(a, (int)(-g)));
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

__device__ __forceinline__ int ndigits(uint32_t v)
{
    int n = 0;
    while (v) { n++; v /= 10; }
    return n;
}

/* Block-wide append: one global atomic per workgroup (same-address atomics
 * from every wave serialise at one L2 channel).  Every thread of the block
 * must reach the call; s_tmp is __shared__ uint32_t[2]. */
__device__ __forceinline__ void block_append(bool pred, uint32_t value, uint32_t *list, uint32_t *count,
                                             uint32_t *s_tmp)
{
    const uint64_t m = __ballot(pred);
    const uint32_t lane = __lane_id();
    if (threadIdx.x == 0) s_tmp[0] = 0;
    __syncthreads();
    uint32_t woff = 0;
    if (lane == 0 && m) woff = atomicAdd(&s_tmp[0], (uint32_t)__popcll(m));
    woff = __shfl(woff, 0);
    __syncthreads();
    if (threadIdx.x == 0) s_tmp[1] = s_tmp[0] ? atomicAdd(count, s_tmp[0]) : 0u;
    __syncthreads();
    if (pred) list[s_tmp[1] + woff + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = value;
}

/* Wave-wide inclusive prefix sum (64 lanes). */
__device__ __forceinline__ uint32_t shfl_up_t(uint32_t v, int d) { return (uint32_t)__shfl_up((int)v, d, 64); }
__device__ __forceinline__ uint64_t shfl_up_t(uint64_t v, int d)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
}

/* The same for u32 in six DPP adds, no LDS: Hillis-Steele within each row
 * of 16 lanes (row_shr 1, 2, 4, 8), then the row totals carried by
 * row_bcast:15 (into rows 1, 3) and row_bcast:31 (into rows 2, 3).  Lanes
 * whose source is out of range (or masked off) add 0. */
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T u = shfl_up_t(v, d);
        if (lane >= (uint32_t)d) v += u;
    }
    return v;
}

/* Inclusive scan over the workgroup (NT threads); s_w holds NT/64 values. */
template <typename T, int NT>
__device__ __forceinline__ T block_incl_scan(T v, T *s_w, T *total)
{
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    v = wave_incl_scan(v);
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
        const T x = s_w[i];
        if ((uint32_t)i < w) off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return v + off;
}

/* ------------------------------------------------------------------ k_seed */

#define SEED_READS 64        /* reads per workgroup (2 threads each) */
#define SEED_ROW 68u         /* LDS bytes per staged read (64 used) */

#define NSEL_HEAVY 0x80000000u
/* some selected 12-mer's bucket is not in ascending order: the item is
 * planned by k_plan_irreg (block_heap_sort restated), see there */
#define NSEL_IRREG 0x40000000u
#define NSEL_FLAGS (NSEL_HEAVY | NSEL_IRREG)

/* rows of the per-length candidate counts: workgroup w adds to row w % 8, so
 * that one read length per batch does not put every workgroup's atomic on
 * one address */
#define LBIN_ROWS 8u

struct SeedOut {
    uint32_t *words;    /* [nrs][32]: 16 interleaved 2-bit words, 8 plane pairs */
    swm_sel_t *sel;     /* [nrs][seln] */
    uint32_t *selc;     /* [nrs][seln]: counts of SEL_BIG or more, written only for those */
    uint32_t *nsel;     /* [nrs] selected 12-mers (4 per selected long seed) */
    uint64_t *npos;     /* [nrs + 1] positions over the selected seeds */
    uint32_t *nchunk;   /* [nrs + 1] sum of ceil(cnt / 32), + 1 */
    unsigned long long *counters;   /* the run's counters, zeroed here */
    uint32_t *nuniq;    /* [nrs + 1]: its last entry is zeroed here */
    uint32_t *nbatch;   /* [nrs]: 1 here, the batch count of an item planned in more (k_plan_pair leaves it) */
    uint32_t *lbins;    /* [LBIN_ROWS][256] candidates per read length, zeroed here */
    const uint32_t *irreg;   /* start offsets of the buckets not in ascending order, ascending */
    uint32_t n_irreg;        /* 0 when the offsets carry the flag instead */
    uint32_t off_mask;       /* the offset bits of a hash_offset entry: 0x7fffffff when bit 31 of a
                              * bucket's start entry flags it irregular, else all ones */
};

/* Is the bucket starting at off one of the n irregular ones (sorted list)? */
__device__ __forceinline__ bool bucket_irregular(const uint32_t *__restrict__ irreg, uint32_t n, uint32_t off)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t v = irreg[mid];
        if (v == off) return true;
        if (v < off) lo = mid + 1; else hi = mid;
    }
    return false;
}

/* A read base as the mapper sees it: A,C,G,T -> 0..3 (complemented on the
 * reverse strand, reverse_read :2429-2450); anything else keeps its byte and
 * is "invalid" (code C in the planes, Q15). */
struct Base {
    uint32_t code;
    bool valid;
};

/* The workgroup's reads are staged in LDS packed: per read 64 bytes of
 * 2-bit base codes (4 bases a byte) and 64 bytes of "not ACGT" flags (4 a
 * byte).  A base as the mapper sees it: A,C,G,T -> 0..3 (complemented on
 * the reverse strand); anything else is invalid and reads as code 1 (C in
 * the planes, Q15). */
__device__ __forceinline__ Base read_base(const uint8_t *sq, const uint8_t *sv, uint32_t L, int strand, uint32_t t)
{
    const uint32_t tt = strand ? L - 1 - t : t;
    const uint32_t c = (sq[tt >> 2] >> (2 * (tt & 3u))) & 3u;
    const bool inv = (sv[tt >> 2] >> (tt & 3u)) & 1u;
    Base b;
    b.valid = !inv;
    b.code = inv ? 1u : (strand ? 3u - c : c);
    return b;
}

/* Seed value of base 15m+u of long seed m (get_seed_from_reads :1260-1300):
 * a non-ACGT base counts as u % 4 when it enters through the k-base loop of
 * the long seed's first 12-mer, and as 3 when it enters as a rolling base. */
__device__ __forceinline__ uint32_t seed_val(const Base &b, uint32_t u, uint32_t k)
{
    return b.valid ? b.code : (u < k ? (u & 3u) : 3u);
}

/* The 12-mer codes of long seed m at shifts q = 0..3 (choose_seed :1372-1390). */
__device__ __forceinline__ void seed_codes(const uint8_t *sq, const uint8_t *sv, uint32_t L, int strand, uint32_t k,
                                           uint32_t base, uint32_t cq[4])
{
    uint32_t code = 0;
    for (uint32_t t = 0; t < k; t++) code |= seed_val(read_base(sq, sv, L, strand, base + t), t, k) << (2 * t);
    cq[0] = code;
#pragma unroll
    for (int q = 1; q < 4; q++) {
        const uint32_t u = k - 1 + (uint32_t)q;
        code = (code >> 2) | (seed_val(read_base(sq, sv, L, strand, base + u), u, k) << (2 * (k - 1)));
        cq[q] = code;
    }
}

/* Packed code byte and invalid nibble of 4 raw bases. */
__device__ __forceinline__ void pack_quad(uint32_t raw, uint8_t *q, uint8_t *v)
{
    uint32_t cb = 0, ib = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t ch = (raw >> (8 * j)) & 0xffu;
        const bool valid = (ch == 'A') | (ch == 'C') | (ch == 'G') | (ch == 'T');
        cb |= (((ch >> 1) ^ (ch >> 2)) & 3u) << (2 * j);     /* A 0, C 1, G 2, T 3 */
        ib |= (valid ? 0u : 1u) << j;
    }
    *q = (uint8_t)cb;
    *v = (uint8_t)ib;
}

/* ------------------------------------------------------------ seed table
 *
 * The offsets and counts of the k-mers (choose_seed's hash_offset gathers,
 * mapper_cpe.c:1372-1390) are kept in a table laid out so that one 128-byte
 * line serves four of a read's lookups.  A long seed's four k-mers c0..c3
 * (shifts q = 0..3) are consecutive: c_q without its first base is c_{q+1}
 * without its last, one shared (k-1)-mer s.  And the reverse strand of a
 * read is its reverse complement, so its k-mers are (for L = 15n exactly,
 * for other lengths in part) the reverse complements of the forward
 * strand's, whose shared (k-1)-mers are rc(s).
 *
 * So every k-mer c is stored twice: in the role "x s" (s = its last k-1
 * bases, x its first base; looked up at even q) and in the role "s y" (s =
 * its first k-1 bases, y its last; odd q), both in the line of the pair
 * {s, rc(s)}: 2 orientations x 2 roles x 4 extension bases = 16 entries of
 * 8 bytes.  The pair (c0, c1) then shares a line, (c2, c3) another, and the
 * reverse strand's (rc(c3), rc(c2)) and (rc(c1), rc(c0)) the same two:
 * the two threads of a read fetch two lines per long seed between them.
 * Lines are indexed by the pair's representative: for odd k-1 the member
 * whose middle base is A or C (reverse complement maps the middle base to
 * its complement), the middle base's high bit dropped -- dense, 4^(k-1) / 2
 * lines; for even k-1 min(s, rc(s)), 4^(k-1) lines of which about half are
 * touched.  Entry (off, count): off with the irregular-bucket flag of the
 * offset entry (bit 31), count the bucket's hits. */
__host__ __device__ __forceinline__ uint32_t rc_bases(uint32_t v, uint32_t nb)
{
    if (nb == 0) return 0u;
    uint32_t r = ~v;
#if defined(__HIP_DEVICE_COMPILE__)
    r = __builtin_bitreverse32(r);
#else
    r = ((r >> 1) & 0x55555555u) | ((r & 0x55555555u) << 1);
    r = ((r >> 2) & 0x33333333u) | ((r & 0x33333333u) << 2);
    r = ((r >> 4) & 0x0f0f0f0fu) | ((r & 0x0f0f0f0fu) << 4);
    r = ((r >> 8) & 0x00ff00ffu) | ((r & 0x00ff00ffu) << 8);
    r = (r >> 16) | (r << 16);
#endif
    /* the bit reversal also swapped the two bits of every base back */
    r = ((r >> 1) & 0x55555555u) | ((r & 0x55555555u) << 1);
    return r >> (32u - 2u * nb);
}

/* the slot of k-mer c in role 0 ("x s") or 1 ("s y") */
__host__ __device__ __forceinline__ uint64_t seed_slot(uint32_t c, uint32_t k, uint32_t role)
{
    const uint32_t n = k - 1u;
    const uint32_t sm = (1u << (2u * n)) - 1u;
    const uint32_t sv = role ? (c & sm) : (c >> 2);
    const uint32_t ext = role ? (c >> (2u * n)) : (c & 3u);
    const uint32_t r = rc_bases(sv, n);
    uint32_t line, o;
    if (n & 1u) {
        const uint32_t h = n >> 1;
        o = ((sv >> (2u * h)) & 3u) >= 2u ? 1u : 0u;
        const uint32_t rep = o ? r : sv;
        const uint32_t lo = rep & ((1u << (2u * h)) - 1u);
        line = lo | (((rep >> (2u * h)) & 1u) << (2u * h)) | ((rep >> (2u * h + 2u)) << (2u * h + 1u));
    } else {
        o = sv > r ? 1u : 0u;
        line = o ? r : sv;
    }
    return (uint64_t)line * 16u + o * 8u + role * 4u + ext;
}

/* the slots of two consecutive k-mers ca (role 0) and cb (role 1), which
 * share their (k-1)-mer: one line, its place computed once */
__host__ __device__ __forceinline__ uint2 seed_pair_slots(uint32_t ca, uint32_t cb, uint32_t k)
{
    const uint64_t a = seed_slot(ca, k, 0u);
    return make_uint2((uint32_t)a, (uint32_t)((a & ~15ull) + (a & 8u) + 4u + (cb >> (2u * (k - 1u)))));
}

/* entries of the table: 8 per (k-1)-mer when k-1 is odd, 16 when even */
__host__ __device__ __forceinline__ uint64_t seed_table_entries(uint32_t k)
{
    return ((k - 1u) & 1u ? 8ull : 16ull) << (2u * (k - 1u));
}

/* both entries of k-mer c: its bucket's start in the device bucket array
 * (with the irregular-bucket flag) and its hit count from the file's offsets */
__global__ void __launch_bounds__(256) k_seed_table(const uint32_t *__restrict__ off, const uint32_t *__restrict__ dev_off,
                                                    uint64_t n_buckets, uint32_t k, uint2 *__restrict__ tab)
{
    const uint64_t c = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (c >= n_buckets) return;
    const uint2 v = make_uint2(dev_off[c], off[c + 1] - off[c]);
    tab[seed_slot((uint32_t)c, k, 0u)] = v;
    tab[seed_slot((uint32_t)c, k, 1u)] = v;
}

/* The device bucket array: bucket c copied from the file's order (src at
 * off[c], n_pos entries) to its place next to its reverse complement's (dst
 * at dev_off[c], n_dev entries), entries past either end left out. */
__global__ void __launch_bounds__(256) k_pl_permute(const uint32_t *__restrict__ off, const uint32_t *__restrict__ dev_off,
                                                    uint64_t n_buckets, uint64_t n_pos, uint64_t n_dev,
                                                    const uint2 *__restrict__ src, uint2 *__restrict__ dst)
{
    const uint64_t c = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (c >= n_buckets) return;
    const uint64_t a = off[c], cnt = (uint32_t)(off[c + 1] - off[c]), d = dev_off[c] & 0x7fffffffu;
    for (uint64_t j = 0; j < cnt && a + j < n_pos && d + j < n_dev; j++) dst[d + j] = src[a + j];
}

/* Work of every seed kernel before its per-thread part: workgroup 0 zeroes
 * the run's counters and the sentinels; the workgroup's reads are staged in
 * LDS packed, one SEED_ROW row each.  Returns the workgroup's read count. */
__device__ __forceinline__ uint32_t seed_stage(const char *__restrict__ reads, uint64_t stride, uint32_t nreads,
                                               SeedOut &o, uint8_t *s_q, uint8_t *s_v)
{
    if (blockIdx.x == 0) {
        /* the run's counters and the sentinel entries the scans read */
        if (threadIdx.x < 12) o.counters[threadIdx.x] = 0;
        if (threadIdx.x == 8) {
            o.npos[2 * nreads] = 0;
            o.nchunk[2 * nreads] = 0;
            o.nuniq[2 * nreads] = 0;
        }
        for (uint32_t j = threadIdx.x; j < LBIN_ROWS * 256u; j += blockDim.x) o.lbins[j] = 0;
    }
    const uint32_t r0 = blockIdx.x * SEED_READS;
    const uint32_t nr = min((uint32_t)SEED_READS, nreads - r0);
    {
        /* quad kq of read lr = its bases 4kq .. 4kq+3 (bytes past the stride,
         * or past the batch, read as 0: they lie beyond every read's length) */
        const char *src = reads + (uint64_t)r0 * stride;
        const uint64_t nbytes = (uint64_t)nr * stride;
        const bool words = (((uintptr_t)src) & 3u) == 0 && (stride & 3u) == 0;
        const uint32_t nq = nr * 64u;
        if (words) {
            /* eight words per thread loaded before any is packed (one load
             * per iteration had put a full memory round trip between the 32
             * words of a thread); a quad past the stride reads the batch's
             * first word and packs as 0 */
            for (uint32_t q0 = threadIdx.x; q0 < nq; q0 += 8u * blockDim.x) {
                uint32_t raw[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const uint32_t qd = q0 + (uint32_t)u * blockDim.x;
                    const uint32_t kq = qd & 63u;
                    const bool ok = qd < nq && 4u * kq < stride;
                    const uint64_t at = ok ? (uint64_t)(qd >> 6) * stride + 4u * kq : 0u;
                    raw[u] = *reinterpret_cast<const uint32_t *>(src + at);
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const uint32_t qd = q0 + (uint32_t)u * blockDim.x;
                    const uint32_t kq = qd & 63u;
                    const uint32_t at = (qd >> 6) * SEED_ROW + kq;
                    if (qd < nq) pack_quad(4u * kq < stride ? raw[u] : 0u, &s_q[at], &s_v[at]);
                }
            }
        }
        /* unaligned reads or stride: byte loads */
        for (uint32_t qd = threadIdx.x; !words && qd < nq; qd += blockDim.x) {
            const uint32_t lr = qd >> 6, kq = qd & 63u;
            uint32_t raw = 0;
            if (4u * kq < stride) {
                const uint64_t at = (uint64_t)lr * stride + 4u * kq;
#pragma unroll
                for (uint32_t j = 0; j < 4; j++)
                    if (4u * kq + j < stride && at + j < nbytes) raw |= (uint32_t)(uint8_t)src[at + j] << (8 * j);
            }
            pack_quad(raw, &s_q[lr * SEED_ROW + kq], &s_v[lr * SEED_ROW + kq]);
        }
    }
    return nr;
}

/* The offset and count of the 12-mers of the first SEED_CACHE_M long seeds
 * of the thread's strand are kept in LDS from the ranking gathers, so the
 * selected ones are not fetched twice; the few selected beyond them (reads
 * of 165 bases and more) and counts of 65535 or more are gathered again.
 * With the reads packed, the workgroup holds 38 KB of LDS, so two waves per
 * SIMD issue gathers: the kernel is bound by the rate of random line fetches
 * (one per lookup), and that rate grows with the waves in flight. */
#define SEED_CACHE_M 10
#ifndef SEED_FIXED_CACHE_M
#define SEED_FIXED_CACHE_M 10
#endif
#define SEED_CHUNK 8

__global__ void __launch_bounds__(128) k_seed(const char *__restrict__ reads, uint64_t stride,
                                              const uint32_t *__restrict__ lens, uint32_t nreads, uint32_t k,
                                              uint32_t e, uint32_t seln, const uint2 *__restrict__ seed_tab,
                                              const swm_recipe_t *__restrict__ recipes, SeedOut o)
{
    /* one row per read, SEED_ROW bytes apart: 17 words, so the rows of a
     * wave's 32 reads start in distinct banks (64-byte rows had put every
     * fourth read's bases in one bank) */
    __shared__ uint8_t s_q[SEED_READS * SEED_ROW], s_v[SEED_READS * SEED_ROW];
    __shared__ uint32_t s_off[4 * SEED_CACHE_M * 2 * SEED_READS];
    __shared__ uint16_t s_cnt[4 * SEED_CACHE_M * 2 * SEED_READS];
    const uint32_t r0 = blockIdx.x * SEED_READS;
    const uint32_t nr = seed_stage(reads, stride, nreads, o, s_q, s_v);
    __syncthreads();
    const uint32_t lr = threadIdx.x >> 1;
    if (lr >= nr) return;
    const int strand = (int)(threadIdx.x & 1u);
    const uint32_t read = r0 + lr;
    const uint32_t r = 2 * read + (uint32_t)strand;
    const uint8_t *sq = s_q + lr * SEED_ROW, *sv = s_v + lr * SEED_ROW;
    uint32_t L = lens[read];
    const uint32_t seed_len = k + SWM_SEED_STEP - 1;
    /* reads that do not fit the reference's fixed tables map to nothing:
     * L + 2 > -l (kseq stride), L > 254 or more than 16 long seeds */
    if ((uint64_t)L + 2 > stride || L > SWM_READ_MAX || L / seed_len > 16) L = 0;
    const uint32_t n_long = L / seed_len;

    /* plain packing + bit planes, then the length's store recipe (Q3) */
    const swm_recipe_t *rec = recipes + L;
    const uint32_t n_planes = rec->n_planes;
    uint32_t pw[18], pl[8], ph[8];
#pragma unroll
    for (int j = 0; j < 8; j++) { pl[j] = 0; ph[j] = 0; }
    pw[16] = pw[17] = 0;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        uint32_t v = 0;
        if ((uint32_t)(16 * j) < L) {
#pragma unroll
            for (int s = 0; s < 16; s++) {
                const uint32_t t = 16 * j + s;
                if (t < L) {
                    Base b = read_base(sq, sv, L, strand, t);
                    v |= b.code << (2 * s);
                    if (t < n_planes) {
                        pl[j >> 1] |= (b.code & 1u) << (t & 31u);
                        ph[j >> 1] |= (b.code >> 1) << (t & 31u);
                    }
                }
            }
        }
        pw[j] = v;
    }
    /* the row in 16-byte stores (the read's LDS row is read again below, so
     * it cannot stage the wave's rows as k_seed_fixed does) */
    uint4 *wo = reinterpret_cast<uint4 *>(o.words + (uint64_t)r * 32);
    {
        uint32_t wv[16];
#pragma unroll
        for (int j = 0; j < 16; j++)
            wv[j] = (pw[j] & rec->mask[j][0]) | (pw[j + 1] & rec->mask[j][1]) | (pw[j + 2] & rec->mask[j][2]);
#pragma unroll
        for (int j = 0; j < 4; j++) wo[j] = make_uint4(wv[4 * j], wv[4 * j + 1], wv[4 * j + 2], wv[4 * j + 3]);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) wo[4 + j] = make_uint4(pl[2 * j], ph[2 * j], pl[2 * j + 1], ph[2 * j + 1]);

    /* 12-mer codes and seed-table gathers, long seed by long seed
     * (choose_seed :1372-1390).  The reverse strand's 12-mer at 15m + q is
     * the reverse complement of the forward strand's at L - 12 - 15m - q,
     * i.e. of long seed n - 1 - m at shift 3 - q when L = 15n (nearby
     * otherwise), so the reverse thread walks its long seeds and shifts in
     * mirrored order: the two threads of a read request each shared line
     * of the seed table in the same load.  Shift q is looked up in role
     * q & 1, so shifts (0, 1) and (2, 3) share a line. */
    uint32_t tot[16], idx[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        tot[m] = 0;
        idx[m] = (uint32_t)m;
    }
    /* SEED_CHUNK long seeds' gathers are issued before any is used: the
     * workgroup's LDS holds two waves per SIMD, so a thread's own loads in
     * flight are what hides the memory latency (one long seed at a time had
     * put a full round trip between every four gathers) */
#pragma unroll
    for (int m0 = 0; m0 < 16; m0 += SEED_CHUNK) {
        if ((uint32_t)m0 < n_long) {
            uint2 a[SEED_CHUNK][4];
#pragma unroll
            for (int u = 0; u < SEED_CHUNK; u++) {
                const int m = m0 + u;
#pragma unroll
                for (int q = 0; q < 4; q++) a[u][q] = make_uint2(0u, 0u);
                if (m < 16 && (uint32_t)m < n_long) {
                    const uint32_t mm = strand ? n_long - 1u - (uint32_t)m : (uint32_t)m;
                    uint32_t cq[4];
                    seed_codes(sq, sv, L, strand, k, mm * seed_len, cq);
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int j = strand ? 3 - q : q;
                        a[u][q] = seed_tab[seed_slot(cq[j], k, (uint32_t)j & 1u)];
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < SEED_CHUNK; u++) {
                const int m = m0 + u;
                if (m < 16 && (uint32_t)m < n_long) {
                    const uint32_t mm = strand ? n_long - 1u - (uint32_t)m : (uint32_t)m;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t c = a[u][q].y;
                        const uint32_t qq = strand ? 3u - (uint32_t)q : (uint32_t)q;
                        tot[m] += c;
                        if (mm < SEED_CACHE_M) {
                            s_off[(4 * mm + qq) * (2 * SEED_READS) + threadIdx.x] = a[u][q].x;    /* with its flag */
                            s_cnt[(4 * mm + qq) * (2 * SEED_READS) + threadIdx.x] = (uint16_t)min(c, 0xffffu);
                        }
                    }
                }
            }
        }
    }
    /* the reverse thread's totals back in long-seed order:
     * tot[j] = walked[n_long - 1 - j] (a static reversal, then a shift by
     * 16 - n_long in four select stages) */
    if (strand) {
        uint32_t rv[16];
#pragma unroll
        for (int j = 0; j < 16; j++) rv[j] = tot[15 - j];
        const uint32_t sh = 16u - n_long;
#pragma unroll
        for (int b = 8; b >= 1; b >>= 1) {
#pragma unroll
            for (int j = 0; j < 16; j++) rv[j] = (sh & (uint32_t)b) ? (j + b < 16 ? rv[j + b] : 0u) : rv[j];
        }
#pragma unroll
        for (int j = 0; j < 16; j++) tot[j] = rv[j];
    }
    /* exchange sort of the long seeds by total hits (:1391-1399, Q4) */
#pragma unroll
    for (int j = 0; j < 16; j++) {
#pragma unroll
        for (int q = j + 1; q < 16; q++) {
            if ((uint32_t)q < n_long && tot[j] > tot[q]) {
                uint32_t t0 = tot[j]; tot[j] = tot[q]; tot[q] = t0;
                uint32_t t1 = idx[j]; idx[j] = idx[q]; idx[q] = t1;
            }
        }
    }
    /* the e+1 rarest long seeds (fewer if the read has fewer, Q5): offsets
     * (cached from the gathers above) and the 3-base residuals of
     * set_read_long_seeds (:2467-2496) */
    const uint32_t n_sel_long = (e + 1 < n_long) ? e + 1 : n_long;
    uint64_t npos = 0;
    uint32_t nchunk = 0;
    bool heavy = false, irr = false;
    swm_sel_t *so = o.sel + (uint64_t)r * seln;
#pragma unroll
    for (int jm = 0; jm < 16; jm++) {
        if ((uint32_t)jm < n_sel_long) {
            const uint32_t m = idx[jm];
            const uint32_t base = m * seed_len;
            uint32_t off4[4], cnt4[4];
            bool again = m >= SEED_CACHE_M;
            if (!again) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    off4[q] = s_off[(4 * m + q) * (2 * SEED_READS) + threadIdx.x];
                    cnt4[q] = s_cnt[(4 * m + q) * (2 * SEED_READS) + threadIdx.x];
                    again |= cnt4[q] == 0xffffu;
                }
            }
            if (again) {
                uint32_t cq[4];
                seed_codes(sq, sv, L, strand, k, base, cq);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint2 v = seed_tab[seed_slot(cq[q], k, (uint32_t)q & 1u)];
                    off4[q] = v.x;
                    cnt4[q] = v.y;
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                uint32_t rl = 0, rh = 0;
#pragma unroll
                for (int bb = 0; bb < 3; bb++) {
                    const uint32_t u = (bb < q) ? (uint32_t)bb : k + (uint32_t)bb;
                    const uint32_t c = read_base(sq, sv, L, strand, base + u).code;
                    rl |= (c & 1u) << bb;
                    rh |= (c >> 1) << bb;
                }
                const uint32_t off = off4[q] & o.off_mask, cnt = cnt4[q];
                const uint32_t meta = (base + (uint32_t)q) | (rl << 8) | (rh << 11) | ((uint32_t)q << 14);
                so[4 * jm + q] = swm_sel_t{off, (min(cnt, SEL_BIG) << 16) | meta};
                if (cnt >= SEL_BIG) o.selc[(uint64_t)r * seln + 4 * jm + q] = cnt;
                heavy |= cnt > SWM_CHUNK;
                if (cnt > 1) {
                    irr |= (off4[q] & ~o.off_mask) != 0u;
                    if (o.n_irreg) irr |= bucket_irregular(o.irreg, o.n_irreg, off);
                }
                npos += cnt;
                nchunk += (cnt + SWM_CHUNK - 1) / SWM_CHUNK;
            }
        }
    }
    /* bit 31: some selected 12-mer has more than one chunk of hits, so the
     * item needs the seed heap (k_plan_heavy) */
    o.nsel[r] = n_sel_long * SWM_SEED_STEP | (heavy ? NSEL_HEAVY : 0u) | (irr ? NSEL_IRREG : 0u);
    o.npos[r] = npos;
    o.nchunk[r] = nchunk + 1;
    o.nbatch[r] = 1;   /* batches <= chunks + 1 (an empty first batch) */
}

/* ---- the seed kernel for a seed length fixed at compile time ----
 *
 * The same results as k_seed, with the read held in registers: each thread
 * loads its read's staged row once (word loads), orients it (the reverse
 * strand's words are the row reversed, complemented and shifted by its
 * length), and takes every base position from compile-time bit offsets --
 * k_seed reads two LDS bytes per base visited (about 900 LDS instructions
 * per wave), which made LDS the kernel's bottleneck. */

/* 2-bit fields of w in reverse order */
__device__ __forceinline__ uint32_t rev_fields(uint32_t w)
{
    const uint32_t r = __builtin_bitreverse32(w);
    return ((r >> 1) & 0x55555555u) | ((r & 0x55555555u) << 1);
}

/* the staged flag word of 16 bases (flags in bits 0-3 of each byte) as a
 * 2-bit mask per base */
__device__ __forceinline__ uint32_t flag_mask2(uint32_t x)
{
    uint32_t f = (x & 0xfu) | ((x >> 4) & 0xf0u) | ((x >> 8) & 0xf00u) | ((x >> 12) & 0xf000u);
    f = (f | (f << 8)) & 0x00ff00ffu;
    f = (f | (f << 4)) & 0x0f0f0f0fu;
    f = (f | (f << 2)) & 0x33333333u;
    f = (f | (f << 1)) & 0x55555555u;
    return f * 3u;
}

/* the even bits of x gathered into its low 16 bits */
__device__ __forceinline__ uint32_t even_bits(uint32_t x)
{
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0f0f0f0fu;
    x = (x | (x >> 4)) & 0x00ff00ffu;
    return (x | (x >> 8)) & 0x0000ffffu;
}

/* a[j] = a[j + sh] (0 past the end) for a per-lane shift sh < 32, in five
 * select stages */
template <int N, typename T>
__device__ __forceinline__ void shift_down(T (&a)[N], uint32_t sh)
{
#pragma unroll
    for (int b = 1; b < N; b <<= 1) {
        const bool on = (sh & (uint32_t)b) != 0u;
#pragma unroll
        for (int j = 0; j < N; j++) a[j] = on ? (j + b < N ? a[j + b] : (T)0) : a[j];
    }
    if (sh >= (uint32_t)N) {
#pragma unroll
        for (int j = 0; j < N; j++) a[j] = (T)0;
    }
}

/* Store half h (words 16 h .. 16 h + 15) of a wave's 64 rows of 32 words,
 * lane l's row at base + 32 l words, through 2 KB of LDS (stage, 16-byte
 * aligned): 32 rows at a time, each row's four 16-byte pieces rotated by a
 * quarter of its row index so that the writing lanes hit distinct banks,
 * then read back in memory order, each store instruction covering 16 rows'
 * halves where a store per lane would touch 64 lines. */
__device__ __forceinline__ void store_half_rows(uint32_t *base, uint32_t h, const uint32_t (&v)[16], uint8_t *stage)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint4 *st4 = reinterpret_cast<uint4 *>(stage);
    uint4 *g4 = reinterpret_cast<uint4 *>(base);
#pragma unroll
    for (uint32_t k = 0; k < 2; k++) {
        if ((lane >> 5) == k) {
            const uint32_t row = lane & 31u;
#pragma unroll
            for (uint32_t j = 0; j < 4; j++)
                st4[row * 4 + ((j + (row >> 2)) & 3u)] = make_uint4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        /* pieces lane and 64 + lane: piece c & 3 of row 32 k + (c >> 2) */
        const uint32_t r0 = lane >> 2, r1 = 16u + (lane >> 2), pc = lane & 3u;
        const uint4 a0 = st4[r0 * 4 + ((pc + (r0 >> 2)) & 3u)];
        const uint4 a1 = st4[r1 * 4 + ((pc + (r1 >> 2)) & 3u)];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        g4[(32 * k + r0) * 8 + 4 * h + pc] = a0;
        g4[(32 * k + r1) * 8 + 4 * h + pc] = a1;
    }
}

/* Store a wave's 64 selection rows of P 16-byte pieces (P <= 16), lane l's
 * row at base + P l pieces, through the same 2 KB of LDS: 8 rows at a time,
 * read back in memory order, each store instruction covering 1 KB. */
__device__ __forceinline__ void store_sel_rows(uint4 *base, uint32_t P, const uint4 (&v)[16], uint8_t *stage)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint4 *st4 = reinterpret_cast<uint4 *>(stage);
#pragma unroll 1
    for (uint32_t k = 0; k < 8; k++) {
        if ((lane >> 3) == k) {
            const uint32_t row = lane & 7u;
#pragma unroll
            for (uint32_t j = 0; j < 16; j++)
                if (j < P) st4[row * P + j] = v[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t n = 8 * P;                            /* pieces of the 8 rows */
        const uint4 a0 = st4[lane < n ? lane : 0u];
        const uint4 a1 = st4[64 + lane < n ? 64 + lane : 0u];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane < n) base[k * n + lane] = a0;
        if (64 + lane < n) base[k * n + 64 + lane] = a1;
    }
}

template <uint32_t K>
__global__ void __launch_bounds__(128) k_seed_fixed(const char *__restrict__ reads, uint64_t stride,
                                                    const uint32_t *__restrict__ lens, uint32_t nreads,
                                                    uint32_t e, uint32_t seln, const uint2 *__restrict__ seed_tab,
                                                    const swm_recipe_t *__restrict__ recipes, SeedOut o)
{
    constexpr uint32_t SL = K + SWM_SEED_STEP - 1;          /* long seed length */
    static_assert(2 * SL <= 32, "a long seed's codes must fit one word");
    constexpr uint32_t SMASK = (uint32_t)((1ull << (2 * SL)) - 1ull);
    constexpr uint32_t KMASK = (1u << (2 * K)) - 1u;
    /* the value of a non-ACGT base at place t of a long seed (seed_val, :1260-1300) */
    constexpr uint32_t SUBST = [] {
        uint32_t v = 0;
        for (uint32_t t = 0; t < SL; t++) v |= (t < K ? (t & 3u) : 3u) << (2 * t);
        return v;
    }();
    __shared__ __align__(16) uint8_t s_q[SEED_READS * SEED_ROW];
    __shared__ __align__(16) uint8_t s_v[SEED_READS * SEED_ROW];
    constexpr uint32_t CM = SEED_FIXED_CACHE_M;
    __shared__ uint32_t s_off[4 * (CM ? CM : 1) * 2 * SEED_READS];
    __shared__ uint16_t s_cnt[4 * (CM ? CM : 1) * 2 * SEED_READS];
    const uint32_t r0 = blockIdx.x * SEED_READS;
    const uint32_t nr = seed_stage(reads, stride, nreads, o, s_q, s_v);
    __syncthreads();
    const uint32_t lr = threadIdx.x >> 1;
    if (lr >= nr) return;
    const int strand = (int)(threadIdx.x & 1u);
    const uint32_t read = r0 + lr;
    const uint32_t r = 2 * read + (uint32_t)strand;
    uint32_t L = lens[read];
    if ((uint64_t)L + 2 > stride || L > SWM_READ_MAX || L / SL > 16) L = 0;
    const uint32_t n_long = L / SL;

    /* the read's codes and non-ACGT masks, oriented to the thread's strand */
    uint32_t cw[16], iv[16];
    {
        const uint32_t *rq = reinterpret_cast<const uint32_t *>(s_q + lr * SEED_ROW);
        const uint32_t *rv = reinterpret_cast<const uint32_t *>(s_v + lr * SEED_ROW);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            cw[j] = rq[j];
            iv[j] = flag_mask2(rv[j]);
        }
    }
    if (strand) {
        /* base t of the reverse strand is the complement of base L - 1 - t:
         * the 512-bit rows reversed field by field, then moved down by
         * 512 - 2L bits */
        uint32_t tc[16], ti[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            tc[j] = rev_fields(~cw[15 - j]);
            ti[j] = __builtin_bitreverse32(iv[15 - j]);
        }
        const uint32_t sh = 512u - 2u * L;
        shift_down(tc, sh >> 5);
        shift_down(ti, sh >> 5);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            cw[j] = __builtin_amdgcn_alignbit(j < 15 ? tc[j + 1] : 0u, tc[j], sh & 31u);
            iv[j] = __builtin_amdgcn_alignbit(j < 15 ? ti[j + 1] : 0u, ti[j], sh & 31u);
        }
    }

    /* plain packing (non-ACGT as 1, Q15; zeros past L), bit planes, the
     * length's store recipe (Q3) */
    const swm_recipe_t *rec = recipes + L;
    const uint32_t n_planes = rec->n_planes;
    uint32_t pw[18];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int rem = (int)L - 16 * j;
        const uint32_t lm = rem >= 16 ? 0xffffffffu : (rem <= 0 ? 0u : (1u << (2 * rem)) - 1u);
        pw[j] = ((cw[j] & ~iv[j]) | (0x55555555u & iv[j])) & lm;
    }
    pw[16] = pw[17] = 0;
    /* the wave's 64 rows are consecutive (r = 2 read + strand): a full wave
     * stages them through its own LDS rows, whose read codes are in
     * registers by now; a partial wave (the last workgroup's) writes its
     * rows directly */
    const bool full_wave = __ballot(1) == ~0ull;
    uint32_t *wrow = o.words + (uint64_t)r * 32;
    uint32_t *wbase = o.words + (uint64_t)(r & ~63u) * 32;
    uint8_t *stage = s_q + (threadIdx.x >> 6) * 32u * SEED_ROW;
    {
        uint32_t wv[16];
#pragma unroll
        for (int j = 0; j < 16; j++)
            wv[j] = (pw[j] & rec->mask[j][0]) | (pw[j + 1] & rec->mask[j][1]) | (pw[j + 2] & rec->mask[j][2]);
        if (full_wave)
            store_half_rows(wbase, 0, wv, stage);
        else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                reinterpret_cast<uint4 *>(wrow)[j] = make_uint4(wv[4 * j], wv[4 * j + 1], wv[4 * j + 2], wv[4 * j + 3]);
        }
    }
    {
        uint32_t wv[16];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int rem = (int)n_planes - 32 * j;
            const uint32_t pm = rem >= 32 ? 0xffffffffu : (rem <= 0 ? 0u : (1u << rem) - 1u);
            const uint32_t lo = pw[2 * j], hi = pw[2 * j + 1];
            wv[2 * j] = (even_bits(lo) | (even_bits(hi) << 16)) & pm;
            wv[2 * j + 1] = (even_bits(lo >> 1) | (even_bits(hi >> 1) << 16)) & pm;
        }
        if (full_wave)
            store_half_rows(wbase, 1, wv, stage);
        else {
#pragma unroll
            for (int j = 0; j < 4; j++)
                reinterpret_cast<uint4 *>(wrow)[4 + j] = make_uint4(wv[4 * j], wv[4 * j + 1], wv[4 * j + 2], wv[4 * j + 3]);
        }
    }

    /* per long seed m: its seed string (the 15 bases as the codes' source,
     * get_seed_from_reads' non-ACGT values) and its residual bits (the
     * bases outside the 12-mer at each shift, set_read_long_seeds
     * :2467-2496, non-ACGT as 1) */
    uint32_t sstr[16], res[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        const uint32_t p2 = 2u * SL * (uint32_t)m;          /* bit offset of the long seed */
        const uint32_t wi = p2 >> 5, bs = p2 & 31u;
        uint32_t raw = 0, inv = 0;
        if (wi < 16) {
            raw = __builtin_amdgcn_alignbit(wi + 1 < 16 ? cw[wi + 1] : 0u, cw[wi], bs) & SMASK;
            inv = __builtin_amdgcn_alignbit(wi + 1 < 16 ? iv[wi + 1] : 0u, iv[wi], bs) & SMASK;
        }
        sstr[m] = (raw & ~inv) | (SUBST & inv);
        const uint32_t code1 = (raw & ~inv) | (0x55555555u & inv);   /* as read_base(...).code */
        const uint32_t plo = even_bits(code1), phi = even_bits(code1 >> 1);
        uint32_t rr = 0;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t keep = (1u << q) - 1u;           /* bases u = bb < q, then K + bb */
            const uint32_t rl = (plo & keep) | ((plo >> K) & 7u & ~keep);
            const uint32_t rh = (phi & keep) | ((phi >> K) & 7u & ~keep);
            rr |= (rl | (rh << 3)) << (6 * q);
        }
        res[m] = rr;
    }
    /* the reverse thread walks its long seeds in mirrored order (see k_seed):
     * walk[m] = sstr[n_long - 1 - m] */
    uint32_t walk[16];
#pragma unroll
    for (int m = 0; m < 16; m++) walk[m] = sstr[m];
    if (strand) {
#pragma unroll
        for (int m = 0; m < 16; m++) walk[m] = sstr[15 - m];
        shift_down(walk, 16u - n_long);
    }

    uint32_t tot[16], pk[16];
#pragma unroll
    for (int m = 0; m < 16; m++) {
        tot[m] = 0;
        pk[m] = (uint32_t)m | (res[m] << 4);                /* seed index | residual bits */
    }
#pragma unroll
    for (int m0 = 0; m0 < 16; m0 += SEED_CHUNK) {
        if ((uint32_t)m0 < n_long) {
            uint2 a[SEED_CHUNK][4];
#pragma unroll
            for (int u = 0; u < SEED_CHUNK; u++) {
                const int m = m0 + u;
#pragma unroll
                for (int q = 0; q < 4; q++) a[u][q] = make_uint2(0u, 0u);
                if (m < 16 && (uint32_t)m < n_long) {
                    /* shifts (0, 1) and (2, 3) share their 11-mer and line */
                    const uint32_t w = walk[m];
                    const uint2 p01 = seed_pair_slots(w & KMASK, (w >> 2) & KMASK, K);
                    const uint2 p23 = seed_pair_slots((w >> 4) & KMASK, (w >> 6) & KMASK, K);
                    const uint32_t sl[4] = {strand ? p23.y : p01.x, strand ? p23.x : p01.y,
                                            strand ? p01.y : p23.x, strand ? p01.x : p23.y};
#pragma unroll
                    for (int q = 0; q < 4; q++) a[u][q] = seed_tab[sl[q]];
                }
            }
#pragma unroll
            for (int u = 0; u < SEED_CHUNK; u++) {
                const int m = m0 + u;
                if (m < 16 && (uint32_t)m < n_long) {
                    const uint32_t mm = strand ? n_long - 1u - (uint32_t)m : (uint32_t)m;
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t c = a[u][q].y;
                        const uint32_t qq = strand ? 3u - (uint32_t)q : (uint32_t)q;
                        tot[m] += c;
                        if (mm < CM) {
                            s_off[(4 * mm + qq) * (2 * SEED_READS) + threadIdx.x] = a[u][q].x;
                            s_cnt[(4 * mm + qq) * (2 * SEED_READS) + threadIdx.x] = (uint16_t)min(c, 0xffffu);
                        }
                    }
                }
            }
        }
    }
    if (strand) {
        uint32_t rv[16];
#pragma unroll
        for (int j = 0; j < 16; j++) rv[j] = tot[15 - j];
        shift_down(rv, 16u - n_long);
#pragma unroll
        for (int j = 0; j < 16; j++) tot[j] = rv[j];
    }
    /* exchange sort of the long seeds by total hits (:1391-1399, Q4), the
     * seed strings and residual bits carried along.  Pass j settles place j
     * and later passes touch only places after it, so the passes past the
     * selected seeds are left out. */
    const uint32_t n_sel_long = (e + 1 < n_long) ? e + 1 : n_long;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        if ((uint32_t)j < n_sel_long) {
#pragma unroll
            for (int q = j + 1; q < 16; q++) {
                if ((uint32_t)q < n_long && tot[j] > tot[q]) {
                    uint32_t t0 = tot[j]; tot[j] = tot[q]; tot[q] = t0;
                    uint32_t t1 = pk[j]; pk[j] = pk[q]; pk[q] = t1;
                    uint32_t t2 = sstr[j]; sstr[j] = sstr[q]; sstr[q] = t2;
                }
            }
        }
    }
    uint64_t npos = 0;
    uint32_t nchunk = 0;
    bool heavy = false, irr = false;
    swm_sel_t *so = o.sel + (uint64_t)r * seln;
    /* rows of at most 32 selections (e <= 7) of a full wave are kept in
     * registers and staged like the words (64 consecutive rows); entries
     * past a read's own selections are written as zeros there */
    const bool stage_sel = full_wave && seln <= 32u;
    uint4 selv[16];
#pragma unroll
    for (int j = 0; j < 16; j++) selv[j] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int jm = 0; jm < 16; jm++) {
        if ((uint32_t)jm < n_sel_long) {
            const uint32_t m = pk[jm] & 15u;
            const uint32_t base = m * SL;
            uint32_t off4[4], cnt4[4];
            bool again = m >= CM;
            if (!again) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    off4[q] = s_off[(4 * m + q) * (2 * SEED_READS) + threadIdx.x];
                    cnt4[q] = s_cnt[(4 * m + q) * (2 * SEED_READS) + threadIdx.x];
                    again |= cnt4[q] == 0xffffu;
                }
            }
            if (again) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint2 v = seed_tab[seed_slot((sstr[jm] >> (2 * q)) & KMASK, K, (uint32_t)q & 1u)];
                    off4[q] = v.x;
                    cnt4[q] = v.y;
                }
            }
            uint32_t sv[8];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t rb = (pk[jm] >> (4 + 6 * q)) & 63u;
                const uint32_t off = off4[q] & o.off_mask, cnt = cnt4[q];
                const uint32_t meta = (base + (uint32_t)q) | (rb << 8) | ((uint32_t)q << 14);
                sv[2 * q] = off;
                sv[2 * q + 1] = (min(cnt, SEL_BIG) << 16) | meta;
                if (cnt >= SEL_BIG) o.selc[(uint64_t)r * seln + 4 * jm + q] = cnt;
                heavy |= cnt > SWM_CHUNK;
                if (cnt > 1) {
                    irr |= (off4[q] & ~o.off_mask) != 0u;
                    if (o.n_irreg) irr |= bucket_irregular(o.irreg, o.n_irreg, off);
                }
                npos += cnt;
                nchunk += (cnt + SWM_CHUNK - 1) / SWM_CHUNK;
            }
            /* a long seed's four selections are 32 bytes at a 32-byte
             * boundary (rows of 4 (e + 1) selections): two 16-byte stores */
            if (jm < 8 && stage_sel) {
                selv[2 * jm] = make_uint4(sv[0], sv[1], sv[2], sv[3]);
                selv[2 * jm + 1] = make_uint4(sv[4], sv[5], sv[6], sv[7]);
            } else {
                uint4 *so4 = reinterpret_cast<uint4 *>(so + 4 * jm);
                so4[0] = make_uint4(sv[0], sv[1], sv[2], sv[3]);
                so4[1] = make_uint4(sv[4], sv[5], sv[6], sv[7]);
            }
        }
    }
    if (stage_sel)
        store_sel_rows(reinterpret_cast<uint4 *>(o.sel + (uint64_t)(r & ~63u) * seln), seln / 2u, selv, stage);
    o.nsel[r] = n_sel_long * SWM_SEED_STEP | (heavy ? NSEL_HEAVY : 0u) | (irr ? NSEL_IRREG : 0u);
    o.npos[r] = npos;
    o.nchunk[r] = nchunk + 1;
    o.nbatch[r] = 1;
}

/* ------------------------------------------------------------ plan kernels */

#define SMALL_BATCH 32
/* Default thresholds of the giant lists (SWM_OPT_GIANT_POS / _CANDS): heavy
 * plan items with more positions, and heavy emission reads with more
 * candidates, are taken first.  The wave priorities below stay keyed to
 * these defaults. */
#define GIANT_POS 16384
#define GIANT_CANDS 16384

struct PlanIO {
    const swm_sel_t *sel;
    const uint32_t *selc;       /* full counts of SEL_BIG or more */
    uint32_t seln;
    const uint32_t *nsel;
    const uint64_t *surv_off;   /* exclusive scan of npos */
    const uint32_t *bd_off;     /* exclusive scan of nchunk */
    uint32_t *surv;             /* survivor / candidate values */
    uint2 *bdesc;               /* per batch: (start relative to surv_off, count) */
    uint32_t *nbatch;           /* [nrs] */
    uint32_t *nuniq;            /* [nrs + 1] */
    uint2 *big;                 /* (rs, batch) of batches left for k_bigsort */
    uint32_t *big_count;
    uint32_t *heavy;            /* (read, strand) items whose seeds need the seed heap */
    uint32_t *heavy_count;
    uint32_t *heavy_head;       /* work-queue head of k_plan_heavy */
    uint32_t *giant;            /* heavy items over the giant threshold (positions): drained first */
    uint32_t *giant_count;
    uint2 *cdesc;               /* heavy items' chunk lists, at bd_off: (bucket offset, batch prefix | seed << 16) */
    struct HBatch *hbatch;      /* heavy items' batches, for k_plan_fetch */
    uint32_t *hb_count;
    const uint32_t *ovf;        /* set when a fixed-capacity run would overflow: every stage stands down */
    uint32_t *irreg;            /* items with an irregular bucket among their selected 12-mers */
    uint32_t *irreg_count;
    uint32_t *merge;            /* heavy batches with a descending chunk (hbatch indices), for k_merge */
    uint32_t *merge_count;
    unsigned long long *light_pos;  /* bucket entries fetched by k_plan_pair */
    uint32_t heavy_cands;       /* k_compact's threshold: items with more candidates read their descriptors */
};

/* Bit 31 of a device bucket entry's second word: its position is smaller
 * than the previous entry's in the same bucket.  A fetched chunk whose
 * entries after the first carry it is not ascending. */
#define PL_DESCENT 0x80000000u

/* One batch of a heavy item: its chunks are cdesc[bd_off[r] + cstart ..
 * + nch), its positions go to surv[surv_off[r] + posoff ..). */
struct HBatch {
    uint32_t r, batch, cstart, npack;   /* npack = nch | pos_size << 8 */
    uint32_t posoff, pad0, pad1, pad2;
};

/* A selection's metadata and count (the full count from the big-count
 * array for SEL_BIG or more; idx = the selection's index in the run). */
__device__ __forceinline__ uint32_t sel_meta(swm_sel_t s) { return s.cm & 0xffffu; }
__device__ __forceinline__ uint32_t sel_cnt(swm_sel_t s, const uint32_t *__restrict__ selc, uint64_t idx)
{
    const uint32_t c = s.cm >> 16;
    return c == SEL_BIG ? selc[idx] : c;
}

/* selection i of item r, decoded */
struct SelV {
    uint32_t off, cnt, meta;
};
__device__ __forceinline__ SelV ld_sel(const PlanIO &io, uint32_t r, uint32_t i)
{
    const uint64_t idx = (uint64_t)r * io.seln + i;
    const swm_sel_t s = io.sel[idx];
    return SelV{s.off, sel_cnt(s, io.selc, idx), sel_meta(s)};
}

/* 15-mer filter of one bucket entry (long_seeds_filter :2106-2117) and the
 * read-start adjustment (adjust_ref_pos_rmdup :1903-1905). */
__device__ __forceinline__ bool filter_pass(uint32_t ls, uint32_t meta)
{
    const uint32_t rl = (meta >> 8) & 7u, rh = (meta >> 11) & 7u, q = (meta >> 14) & 3u;
    return ((((ls >> (3 - q)) ^ rl) | ((ls >> (9 - q)) ^ rh)) & 7u) == 0;
}
__device__ __forceinline__ uint32_t adjust(uint32_t p, uint32_t meta)
{
    const uint32_t bo = meta & 0xffu;
    return p > bo ? p - bo : 0u;
}

/* The plain form of the same state machine: one thread per (read, strand)
 * for every item, the seed heap in registers (no work queue).  Kept as the
 * one-launch-per-stage baseline the work queue is measured against. */
__global__ void __launch_bounds__(64) k_plan_plain(uint32_t nrs, const uint2 *__restrict__ pl, PlanIO io)
{
    if (*io.ovf) return;
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrs) return;
    if (io.nsel[r] & NSEL_IRREG) return;       /* k_plan_irreg's */
    const uint32_t nsel = io.nsel[r] & ~NSEL_FLAGS;
    uint32_t *surv = io.surv + io.surv_off[r];
    uint2 *bd = io.bdesc + io.bd_off[r];

    uint32_t pcount[64];
    uint32_t hloc[66], hind[66];      /* seed heap, 1-based (mapper_cpe.c:2149) */
    uint32_t hsize = 0;
    uint32_t nsurv = 0, bstart = 0, batch = 0, pos_size = 0, blocks = 0;
    bool first_pending = true;        /* Q7 */

    auto take = [&](uint32_t g, uint32_t len, uint32_t meta) {
        uint32_t lv[SWM_CHUNK], pv[SWM_CHUNK];
#pragma unroll
        for (int j = 0; j < SWM_CHUNK; j++) {
            if ((uint32_t)j < len) { const uint2 v = pl[g + j]; pv[j] = v.x; lv[j] = v.y; }
        }
#pragma unroll
        for (int j = 0; j < SWM_CHUNK; j++) {
            if ((uint32_t)j < len) {
                if (first_pending) first_pending = false;
                else if (filter_pass(lv[j], meta)) surv[nsurv++] = adjust(pv[j], meta);
            }
        }
    };
    auto heap_insert = [&](uint32_t loc, uint32_t ind) {      /* :1919-1935 */
        hsize++;
        hloc[hsize] = loc; hind[hsize] = ind;
        uint32_t i = hsize;
        while (i > 1) {
            const uint32_t ti = i / 2;
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

    bool heavy = false;
    for (uint32_t i = 0; i < nsel; i++) {                     /* phase 1 (:2194-2236) */
        const SelV s = ld_sel(io, r, i);
        const uint32_t get = s.cnt > SWM_CHUNK ? SWM_CHUNK : s.cnt;
        if (get > 0) { take(s.off, get, s.meta); blocks++; }
        pos_size += get;
        pcount[i] = get;
        heavy |= get != s.cnt;
    }
    if (heavy) {
        for (uint32_t i = 0; i < nsel; i++) {                 /* :2244-2252 */
            const SelV s = ld_sel(io, r, i);
            if (pcount[i] != s.cnt) heap_insert(pl[s.off + pcount[i] - 1].x, i);
        }
    }
    for (;;) {                                                /* :2261-2419 */
        if (hsize > 0 && pos_size + SWM_CHUNK <= SWM_BATCH_POS && blocks < SWM_MAX_BLOCKS) {
            const uint32_t i = hind[1];
            const SelV s = ld_sel(io, r, i);
            const uint32_t rem = s.cnt - pcount[i];
            const uint32_t get = rem > SWM_CHUNK ? SWM_CHUNK : rem;
            take(s.off + pcount[i], get, s.meta);
            pos_size += get;
            pcount[i] += get;
            blocks++;
            heap_pop_insert(pl[s.off + pcount[i] - 1].x, pcount[i] == s.cnt);
        }
        if (pos_size + SWM_CHUNK > SWM_BATCH_POS || hsize == 0 || blocks >= SWM_MAX_BLOCKS) {
            const uint32_t n = nsurv - bstart;
            if (n <= SMALL_BATCH) {
                uint32_t *a = surv + bstart;
                for (uint32_t i = 1; i < n; i++) {
                    uint32_t v = a[i], j = i;
                    while (j > 0 && a[j - 1] > v) { a[j] = a[j - 1]; j--; }
                    a[j] = v;
                }
                uint32_t u = n ? 1u : 0u;
                for (uint32_t i = 1; i < n; i++)
                    if (a[i] != a[u - 1]) a[u++] = a[i];
                bd[batch] = make_uint2(bstart, u);
                nsurv = bstart + u;
            } else {
                bd[batch] = make_uint2(bstart, n);
                const uint32_t slot = atomicAdd(io.big_count, 1u);
                io.big[slot] = make_uint2(r, batch);
            }
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

/* block_heap_sort (mapper_cpe.c:1999-2089) restated for one batch: a k-way
 * merge of the batch's blocks (one per fetch, each the chunk's surviving
 * read starts in fetch order) through a binary heap keyed by each block's
 * current head, an output skipped when it equals the one before.  On
 * ascending blocks this is the sorted unique list the other plan kernels
 * build; on a block that descends (a bucket of the reference's index that
 * its serial scatter left out of order: the zeroed tail starts of Q23 land
 * at the head of a later bucket, and their own bucket keeps an unwritten 0
 * slot) the merge's order and its surviving duplicates depend on the heap's
 * steps, so they are taken literally.  pos: the blocks back to back; ind:
 * [2b] start, [2b + 1] end of block b (advanced in place); returns the
 * count written to out.  An empty batch gives nothing (Q8). */
__device__ uint32_t block_merge(const uint32_t *pos, uint32_t *ind, uint32_t nblk, uint32_t *hp, uint32_t *hb,
                                uint32_t *__restrict__ out)
{
    uint32_t hs = 0;
    for (uint32_t i = 0; i < nblk; i++) {                   /* :2003-2026 */
        if (ind[2 * i] < ind[2 * i + 1]) {
            hs++;
            hp[hs] = pos[ind[2 * i]];
            hb[hs] = 2 * i;
            ind[2 * i]++;
            uint32_t cur = hs;
            while (cur > 1) {
                const uint32_t par = cur / 2;
                if (hp[par] > hp[cur]) {
                    uint32_t t = hp[cur]; hp[cur] = hp[par]; hp[par] = t;
                    t = hb[cur]; hb[cur] = hb[par]; hb[par] = t;
                } else {
                    break;
                }
                cur = par;
            }
        }
    }
    if (hs == 0) return 0;
    uint32_t pre = 0xffffffffu, cnt = 0;
    while (hs > 1) {                                        /* :2031-2074 */
        const uint32_t id = hb[1];
        if (pre != hp[1]) { out[cnt++] = hp[1]; pre = hp[1]; }
        if (ind[id] < ind[id + 1]) {
            hp[1] = pos[ind[id]];
            ind[id]++;
        } else {
            hp[1] = hp[hs];
            hb[1] = hb[hs];
            hs--;
            if (hs == 1) break;
        }
        uint32_t h = 1;
        while (h * 2 <= hs) {
            const uint32_t l = h * 2, rr = l + 1;
            uint32_t sm = l;
            if (rr <= hs && hp[l] > hp[rr]) sm = rr;
            if (hp[h] > hp[sm]) {
                uint32_t t = hp[h]; hp[h] = hp[sm]; hp[sm] = t;
                t = hb[h]; hb[h] = hb[sm]; hb[sm] = t;
            } else {
                break;
            }
            h = sm;
        }
    }
    const uint32_t id = hb[1];                              /* :2077-2084 */
    for (uint32_t i = ind[id] - 1; i < ind[id + 1]; i++) {
        const uint32_t v = pos[i];
        if (v != pre) { out[cnt++] = v; pre = v; }
    }
    return cnt;
}

/* Items whose selected 12-mers include an irregular bucket (NSEL_IRREG):
 * the whole alignment_rmdup state machine (:2124-2426) for the item on one
 * lane, as k_plan_plain, with each batch's blocks kept in LDS and merged by
 * block_merge.  Such items are rare (the reference's index has at most a
 * few irregular buckets), so one lane per item, a workgroup each, taken
 * from the list grid-stride; the item's plan output (batch descriptors,
 * candidates, counts) has the layout every other plan kernel writes. */
__global__ void __launch_bounds__(64) k_plan_irreg(const uint2 *__restrict__ pl, PlanIO io)
{
    __shared__ uint32_t s_pos[SWM_BATCH_POS];
    __shared__ uint32_t s_ind[2 * SWM_MAX_BLOCKS + 4];
    __shared__ uint32_t s_hp[SWM_MAX_BLOCKS + 2], s_hb[SWM_MAX_BLOCKS + 2];
    __shared__ uint32_t s_loc[68], s_sid[68], s_pc[64];
    __shared__ uint2 s_stage[64 * SWM_CHUNK];      /* each 12-mer's first chunk, staged by the wave */
    if (*io.ovf) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t nq = *io.irreg_count;
    for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const uint32_t r = io.irreg[qi];
        const uint32_t nsel = io.nsel[r] & ~NSEL_FLAGS;
        /* the first chunk of every selected 12-mer (all of phase 1's
         * fetches, and the heap keys of the longer ones) fetched by the whole
         * wave, eight 12-mers' loads in flight at once: lane 0 then replays
         * the state machine from LDS, where one fetch per chunk had put a
         * memory round trip on its serial path for each 12-mer */
        {
            const SelV mine = lane < nsel ? ld_sel(io, r, lane) : SelV{0, 0, 0};
            for (uint32_t i0 = 0; i0 < nsel; i0 += 8) {
                uint2 v[8];
#pragma unroll
                for (uint32_t u = 0; u < 8; u++) {
                    const uint32_t i = i0 + u;
                    v[u] = make_uint2(0u, 0u);
                    if (i < nsel) {
                        const uint32_t cnt = (uint32_t)__builtin_amdgcn_readlane((int)mine.cnt, (int)i), off = (uint32_t)__builtin_amdgcn_readlane((int)mine.off, (int)i);
                        if (lane < min(cnt, (uint32_t)SWM_CHUNK)) v[u] = pl[off + lane];
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < 8; u++) {
                    const uint32_t i = i0 + u;
                    if (i < nsel && lane < min((uint32_t)__builtin_amdgcn_readlane((int)mine.cnt, (int)i), (uint32_t)SWM_CHUNK)) s_stage[i * SWM_CHUNK + lane] = v[u];
                }
            }
        }
        __syncthreads();
        if (lane == 0) {
            uint32_t *surv = io.surv + io.surv_off[r];
            uint2 *bd = io.bdesc + io.bd_off[r];
            uint32_t hsize = 0, nsurv = 0, batch = 0, pos_size = 0, blocks = 0, nb = 0;
            bool first_pending = true;                          /* Q7 */
            auto take = [&](uint32_t g, uint32_t len, uint32_t meta) {
                const uint32_t b0 = nb;
                uint2 v[SWM_CHUNK];
#pragma unroll
                for (int j = 0; j < SWM_CHUNK; j++) v[j] = pl[(uint32_t)j < len ? g + (uint32_t)j : g];   /* all in flight */
#pragma unroll
                for (int j = 0; j < SWM_CHUNK; j++) {
                    if ((uint32_t)j < len) {
                        if (first_pending) first_pending = false;
                        else if (filter_pass(v[j].y, meta)) s_pos[nb++] = adjust(v[j].x, meta);
                    }
                }
                s_ind[2 * blocks] = b0;
                s_ind[2 * blocks + 1] = nb;
                blocks++;
            };
            auto heap_insert = [&](uint32_t loc, uint32_t ind) {    /* :1919-1935 */
                hsize++;
                s_loc[hsize] = loc; s_sid[hsize] = ind;
                uint32_t i = hsize;
                while (i > 1) {
                    const uint32_t ti = i / 2;
                    if (s_loc[ti] > s_loc[i]) {
                        uint32_t t = s_loc[ti]; s_loc[ti] = s_loc[i]; s_loc[i] = t;
                        t = s_sid[ti]; s_sid[ti] = s_sid[i]; s_sid[i] = t;
                    } else break;
                    i = ti;
                }
            };
            auto heap_pop_insert = [&](uint32_t loc, bool del) {    /* :1936-1962 (Q6) */
                if (del) { s_loc[1] = s_loc[hsize]; s_sid[1] = s_sid[hsize]; hsize--; }
                else s_loc[1] = loc;
                uint32_t i = 1;
                while (i * 2 < hsize) {
                    uint32_t ti = i * 2;
                    if (i * 2 + 1 <= hsize && s_loc[i * 2] > s_loc[i * 2 + 1]) ti = i * 2 + 1;
                    if (s_loc[ti] < s_loc[i]) {
                        uint32_t t = s_loc[ti]; s_loc[ti] = s_loc[i]; s_loc[i] = t;
                        t = s_sid[ti]; s_sid[ti] = s_sid[i]; s_sid[i] = t;
                    } else break;
                    i = ti;
                }
            };
            for (uint32_t i = 0; i < nsel; i++) {                   /* phase 1 (:2194-2236), from LDS */
                const SelV sv = ld_sel(io, r, i);
                const uint32_t get = sv.cnt > SWM_CHUNK ? SWM_CHUNK : sv.cnt;
                if (get > 0) {
                    const uint32_t b0 = nb;
                    for (uint32_t j = 0; j < get; j++) {
                        const uint2 v = s_stage[i * SWM_CHUNK + j];
                        if (first_pending) first_pending = false;
                        else if (filter_pass(v.y, sv.meta)) s_pos[nb++] = adjust(v.x, sv.meta);
                    }
                    s_ind[2 * blocks] = b0;
                    s_ind[2 * blocks + 1] = nb;
                    blocks++;
                }
                pos_size += get;
                s_pc[i] = get;
            }
            for (uint32_t i = 0; i < nsel; i++) {                   /* :2244-2252 */
                const SelV sv = ld_sel(io, r, i);
                if (s_pc[i] != sv.cnt) heap_insert(s_stage[i * SWM_CHUNK + s_pc[i] - 1].x, i);
            }
            for (;;) {                                              /* :2261-2419 */
                if (hsize > 0 && pos_size + SWM_CHUNK <= SWM_BATCH_POS && blocks < SWM_MAX_BLOCKS) {
                    const uint32_t i = s_sid[1];
                    const SelV sv = ld_sel(io, r, i);
                    const uint32_t rem = sv.cnt - s_pc[i];
                    const uint32_t get = rem > SWM_CHUNK ? SWM_CHUNK : rem;
                    take(sv.off + s_pc[i], get, sv.meta);
                    pos_size += get;
                    s_pc[i] += get;
                    heap_pop_insert(pl[sv.off + s_pc[i] - 1].x, s_pc[i] == sv.cnt);
                }
                if (pos_size + SWM_CHUNK > SWM_BATCH_POS || hsize == 0 || blocks >= SWM_MAX_BLOCKS) {
                    const uint32_t u = block_merge(s_pos, s_ind, blocks, s_hp, s_hb, surv + nsurv);
                    bd[batch] = make_uint2(nsurv, u);
                    nsurv += u;
                    batch++;
                    pos_size = 0;
                    blocks = 0;
                    nb = 0;
                    first_pending = true;
                    if (hsize == 0) break;
                }
            }
            io.nbatch[r] = batch;
            io.nuniq[r] = nsurv;
        }
        __syncthreads();
    }
}
/* Lane reads (v_readlane) and writes (a compare + select) at a wave-uniform
 * lane index.  The seed heap and the per-seed state of k_plan_heavy live one
 * entry per lane, so its replay is scalar code with no LDS round trip per
 * heap level. */
__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ uint32_t wl(uint32_t v, uint32_t l, uint32_t x)
{
    return __lane_id() == l ? x : v;
}
/* a > b ? x : y on scalars, as s_cmp + s_cselect (the compiler's own form
 * of this select goes through a vector register and back) */
__device__ __forceinline__ uint32_t sel_gt(uint32_t a, uint32_t b, uint32_t x, uint32_t y)
{
    uint32_t r;
    asm("s_cmp_gt_u32 %1, %2\n\ts_cselect_b32 %0, %3, %4" : "=s"(r) : "s"(a), "s"(b), "s"(x), "s"(y) : "scc");
    return r;
}

/* Light items (every selected 12-mer with at most one chunk of hits), both
 * strands of a read in one wave.  A strand's selected 12-mers' entries are
 * laid end to end (a prefix sum of the counts) and fetched 64 at a time,
 * lane j on entry e0 + j, each lane finding its 12-mer by binary search
 * over the prefix sums: the lanes of one round read consecutive addresses
 * and one line costs one request.  W = 32 (e <= 7): lanes 0-31 hold the
 * forward strand's selections and lanes 32-63 the reverse one's, so one
 * prefix sum lays both strands' entries end to end and one batch of loads
 * fetches them; W = 64 (any e): each lane holds both strands' selection of
 * its index, two scans.  The strands select mostly the same long seeds,
 * whose 12-mers are reverse complements of each other, and each bucket lies
 * beside its reverse complement's, so the two strands' requests for a line
 * reach the L2 together and cost it one memory read -- fetched one strand
 * after the other they cost two (the line is gone from the L2 by then).
 * Results as k_plan: per strand, the 15-mer filter and read-start
 * adjustment in entry order with the first entry dropped (Q7), then the
 * sorted unique list of the one batch (<= 64 survivors sorted in the wave
 * by rank, larger ones left to k_bigsort).  Waves take reads one at a time
 * (grid-stride), the next read's selections loaded while this one's are
 * processed; a strand that is heavy or irregular is left to its kernels. */
#define PP_ROUNDS 3

template <uint32_t W>   /* selection slots per strand: 32 (e <= 7) or 64 (any e) */
__global__ void __launch_bounds__(64) k_plan_pair(uint32_t nreads, const uint2 *__restrict__ pl, PlanIO io)
{
    /* slots: W = 32, slot s = lane (lanes 0-31 forward, 32-63 reverse);
     * W = 64, slot lane holds the forward selection `lane` and slot 64 + lane
     * the reverse one's */
    constexpr uint32_t NS = 2 * W;
    __shared__ uint32_t s_off[NS], s_pre[NS], s_meta[NS], s_surv[2][64];
    if (*io.ovf) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t half = lane >> 5, li = lane & 31u;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t fetched = 0;
    uint32_t nsw_n = 0, nsw_nb = 0;
    swm_sel_t my_n = {0, 0}, my_nb = {0, 0};
    uint64_t so_n0 = 0, so_n1 = 0;
    uint32_t bo_n0 = 0, bo_n1 = 0;
    /* the next read's selection counts, selections and output offsets are
     * loaded while this one is processed: the descriptor offsets otherwise
     * put a round trip between the survivors' dedup and its stores */
    auto prefetch = [&](uint32_t rd) {
        so_n0 = io.surv_off[2u * rd];
        so_n1 = io.surv_off[2u * rd + 1u];
        bo_n0 = io.bd_off[2u * rd];
        bo_n1 = io.bd_off[2u * rd + 1u];
        if (W == 32) {
            const uint32_t r = 2u * rd + half;
            nsw_n = io.nsel[r];
            my_n = li < io.seln ? io.sel[(uint64_t)r * io.seln + li] : swm_sel_t{0, 0};
        } else {
            nsw_n = io.nsel[2u * rd];
            nsw_nb = io.nsel[2u * rd + 1u];
            my_n = lane < io.seln ? io.sel[(uint64_t)(2u * rd) * io.seln + lane] : swm_sel_t{0, 0};
            my_nb = lane < io.seln ? io.sel[(uint64_t)(2u * rd + 1u) * io.seln + lane] : swm_sel_t{0, 0};
        }
    };
    if (blockIdx.x < nreads) prefetch(blockIdx.x);
    for (uint32_t rd = blockIdx.x; rd < nreads; rd += gridDim.x) {
        const uint32_t nsw = nsw_n, nswb = nsw_nb;
        swm_sel_t my = my_n, myb = my_nb;
        const uint64_t so0 = so_n0, so1 = so_n1;
        const uint32_t bo0 = bo_n0, bo1 = bo_n1;
        if (rd + gridDim.x < nreads) prefetch(rd + gridDim.x);
        bool light0, light1;
        uint32_t total, total0;
        if (W == 32) {
            const bool mine = (nsw & NSEL_FLAGS) == 0;      /* heavy / irregular: another kernel's */
            light0 = rl(mine ? 1u : 0u, 0) != 0u;
            light1 = rl(mine ? 1u : 0u, 32) != 0u;
            if (!light0 && !light1) continue;
            if (!mine || li >= nsw) my = swm_sel_t{0, 0};
            const uint32_t cnt = my.cm >> 16;               /* light: every count is at most a chunk */
            const uint32_t incl = wave_incl_scan_dpp(cnt);
            total = rl(incl, 63);
            total0 = rl(incl, 31);
            s_off[lane] = my.off;
            s_pre[lane] = incl - cnt;
            s_meta[lane] = sel_meta(my);
        } else {
            light0 = (nsw & NSEL_FLAGS) == 0;
            light1 = (nswb & NSEL_FLAGS) == 0;
            if (!light0 && !light1) continue;
            if (!light0 || lane >= nsw) my = swm_sel_t{0, 0};
            if (!light1 || lane >= nswb) myb = swm_sel_t{0, 0};
            const uint32_t cnt = my.cm >> 16, cntb = myb.cm >> 16;
            const uint32_t incl = wave_incl_scan_dpp(cnt), inclb = wave_incl_scan_dpp(cntb);
            total0 = rl(incl, 63);
            total = total0 + rl(inclb, 63);
            s_off[lane] = my.off;
            s_pre[lane] = incl - cnt;
            s_meta[lane] = sel_meta(my);
            s_off[64 + lane] = myb.off;
            s_pre[64 + lane] = total0 + inclb - cntb;
            s_meta[64 + lane] = sel_meta(myb);
        }
        fetched += total;
        __syncthreads();
        uint32_t *surv0 = io.surv + so0, *surv1 = io.surv + so1;
        uint32_t ns0 = 0, ns1 = 0;
        for (uint32_t e0 = 0; e0 < total; e0 += 64u * PP_ROUNDS) {
            uint32_t lo[PP_ROUNDS], meta[PP_ROUNDS];
            uint2 v[PP_ROUNDS];
#pragma unroll
            for (int j = 0; j < PP_ROUNDS; j++) lo[j] = 0;
#pragma unroll
            for (uint32_t hb = NS / 2; hb >= 1; hb >>= 1) {
#pragma unroll
                for (int j = 0; j < PP_ROUNDS; j++) {
                    const uint32_t ent = e0 + 64u * (uint32_t)j + lane;
                    if (s_pre[lo[j] + hb] <= ent) lo[j] += hb;
                }
            }
#pragma unroll
            for (int j = 0; j < PP_ROUNDS; j++) {
                const uint32_t ent = e0 + 64u * (uint32_t)j + lane;
                meta[j] = s_meta[lo[j]];
                v[j] = make_uint2(0u, 0u);
                if (e0 + 64u * (uint32_t)j < total)       /* wave-uniform */
                    v[j] = pl[ent < total ? s_off[lo[j]] + (ent - s_pre[lo[j]]) : s_off[0]];
            }
#pragma unroll
            for (int j = 0; j < PP_ROUNDS; j++) {
                const uint32_t ent = e0 + 64u * (uint32_t)j + lane;
                const bool st1 = ent >= total0;
                /* each strand's first fetched entry is dropped (Q7) */
                const bool keep = ent < total && ent != (st1 ? total0 : 0u) && filter_pass(v[j].y, meta[j]);
                const uint64_t m = __ballot(keep), m1 = __ballot(keep && st1), m0 = m & ~m1;
                if (keep) {
                    const uint32_t at = st1 ? ns1 + (uint32_t)__popcll(m1 & below) : ns0 + (uint32_t)__popcll(m0 & below);
                    const uint32_t val = adjust(v[j].x, meta[j]);
                    /* the first 64 of a strand stay in LDS: they reach memory
                     * only as uniqued values, or below if the batch is large */
                    if (at < 64u) s_surv[st1 ? 1 : 0][at] = val;
                    else (st1 ? surv1 : surv0)[at] = val;
                }
                ns0 += (uint32_t)__popcll(m0);
                ns1 += (uint32_t)__popcll(m1);
            }
        }
        __syncthreads();
        uint32_t nu[2] = {0u, 0u};
#pragma unroll
        for (uint32_t h = 0; h < 2; h++) {
            if (!(h ? light1 : light0)) continue;
            const uint32_t r = 2u * rd + h;
            const uint32_t nsurv = h ? ns1 : ns0;
            uint32_t *surv = h ? surv1 : surv0;
            uint2 *bd = io.bdesc + (h ? bo1 : bo0);
            if (nsurv <= 64u) {
                const bool have = lane < nsurv;
                const uint32_t v = have ? s_surv[h][lane] : 0xffffffffu;
                uint64_t fm = 0;
                for (uint64_t left = __ballot(have); left;) {
                    const uint32_t f = (uint32_t)__ffsll((unsigned long long)left) - 1u;
                    const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)v, (int)f);
                    fm |= 1ull << f;
                    left &= ~__ballot(have && v == w);
                }
                uint32_t rank = 0;
                for (uint64_t f = fm; f; f &= f - 1ull) {
                    const uint32_t q = (uint32_t)__ffsll((unsigned long long)f) - 1u;
                    rank += (uint32_t)__builtin_amdgcn_readlane((int)v, (int)q) < v ? 1u : 0u;
                }
                if ((fm >> lane) & 1ull) surv[rank] = v;
                const uint32_t u = (uint32_t)__popcll(fm);
                /* a one-batch item's descriptor is read only by
                 * k_compact_heavy (k_compact takes it from the count) */
                if (lane == 0 && u > io.heavy_cands) bd[0] = make_uint2(0, u);
                nu[h] = u;
            } else {
                surv[lane] = s_surv[h][lane];
                if (lane == 0) {
                    bd[0] = make_uint2(0, nsurv);
                    io.big[atomicAdd(io.big_count, 1u)] = make_uint2(r, 0);
                }
                nu[h] = nsurv;
            }
        }
        /* both strands' counts in one two-lane store (their batch counts,
         * 1, were written by the seed kernel) */
        if (lane < 2u && (lane ? light1 : light0)) io.nuniq[2u * rd + lane] = lane ? nu[1] : nu[0];
        __syncthreads();
    }
    if (lane == 0 && fetched) atomicAdd(io.light_pos, (unsigned long long)fetched);
}

/* Heavy (read, strand): a device work queue drained by persistent
 * single-wave workgroups.  alignment_rmdup's batch composition -- which
 * 32-position chunks, in what heap order, where each batch ends -- depends
 * only on the selected seeds' counts and on the raw last position of every
 * fetched chunk (the heap key, Q6), never on the filter.  So k_plan_heavy
 * only replays the seed heap (:1919-1962, 2261-2356) from keys prefetched
 * into LDS and writes each batch's chunk list; the batches' positions are
 * fetched, filtered and compacted by k_plan_fetch, all batches in parallel,
 * each at its own offset (its first position's index in the item), so the
 * serial part of a giant item is the heap replay alone.
 *
 * The replay is one dependent chain per item (a pop decides the next root),
 * so a pop is kept short: the heap keys of each 12-mer's next two chunks sit
 * in registers (one lane per 12-mer) and the key after them is read from an
 * LDS ring during the sift-down and written one pop later; the sift-down is
 * unrolled (scalar selects on lane reads); the ring is
 * refilled only when some 12-mer runs low, and then only its consumed
 * slots are loaded. */
#define HK 16          /* ring of buffered chunk keys per selected 12-mer */
#define HK_LOW 6       /* refill when a 12-mer has fewer buffered keys left */

__global__ void __launch_bounds__(64) k_plan_heavy(const uint2 *__restrict__ pl, PlanIO io)
{
    __shared__ uint32_t skey[64 * HK];           /* key of chunk c of 12-mer i at [i * HK + c % HK] */
    __shared__ uint32_t sv[1];
    /* the heavy lists are filled whatever the capacities: on overflow the
     * positions and chunks they would write lie beyond their buffers */
    if (*io.ovf) return;
    const uint32_t lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t ng = *io.giant_count, nq = ng + *io.heavy_count;
    /* batch records wait in registers (record j in lane j) and are appended
     * 64 at a time: one atomic on the shared counter per 64 batches */
    uint32_t hq_r = 0, hq_b = 0, hq_c = 0, hq_n = 0, hq_p = 0, nhq = 0;
    auto hb_flush = [&]() {
        uint32_t at = 0;
        if (lane == 0 && nhq) at = atomicAdd(io.hb_count, nhq);
        at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
        if (lane < nhq) {
            HBatch hb;
            hb.r = hq_r; hb.batch = hq_b; hb.cstart = hq_c; hb.npack = hq_n;
            hb.posoff = hq_p; hb.pad0 = hb.pad1 = hb.pad2 = 0;
            io.hbatch[at + lane] = hb;
        }
        nhq = 0;
    };

    /* a wave's first item is its workgroup index, later ones come from the
     * queue head (offset by the grid): no atomic for waves that find no work */
    for (bool first = true;; first = false) {
        if (lane == 0) sv[0] = first ? blockIdx.x : gridDim.x + atomicAdd(io.heavy_head, 1u);
        __syncthreads();
        const uint32_t qi = sv[0];
        __syncthreads();
        if (qi >= nq) break;
        const uint32_t r = qi < ng ? io.giant[qi] : io.heavy[qi - ng];
        {
            /* the longest replays set the kernel's length: they issue first */
            const uint64_t np = io.surv_off[r + 1] - io.surv_off[r];
            if (np > 16u * GIANT_POS) __builtin_amdgcn_s_setprio(3);
            else if (np > 4u * GIANT_POS) __builtin_amdgcn_s_setprio(2);
            else if (np > GIANT_POS) __builtin_amdgcn_s_setprio(1);
            else __builtin_amdgcn_s_setprio(0);
        }
        const uint32_t nsel = io.nsel[r] & ~NSEL_FLAGS;
        const bool act = lane < nsel;
        /* per 12-mer state, lane i = selected 12-mer i */
        SelV my = {0, 0, 0};
        if (act) my = ld_sel(io, r, lane);
        const uint32_t cnt_v = my.cnt, off_v = my.off;
        uint32_t spc_v = min(cnt_v, (uint32_t)SWM_CHUNK);     /* positions fetched */
        /* heap key after fetching chunk c: the raw last position of that
         * chunk (seeds_heap_pop_and_insert's pos argument, :2350).  The
         * ring holds chunks kb .. kb + HK - 1; (re)fill chunks
         * max(lo, kbn) .. kbn + HK - 1 and make kbn the new base */
        uint32_t kb_v = 1;
        auto fill = [&](bool live, uint32_t kbn, uint32_t lo) {
            /* all loads first (clamped to an entry of the item), then the stores */
            uint32_t kv[HK];
#pragma unroll
            for (uint32_t k = 0; k < HK; k++) {
                const uint32_t c = kbn + k;
                const bool ok = live && c >= lo && SWM_CHUNK * c < cnt_v;
                kv[k] = pl[ok ? off_v + min(SWM_CHUNK * (c + 1), cnt_v) - 1 : off_v].x;
            }
#pragma unroll
            for (uint32_t k = 0; k < HK; k++) {
                const uint32_t c = kbn + k;
                if (live && c >= lo && SWM_CHUNK * c < cnt_v) skey[lane * HK + (c & (HK - 1))] = kv[k];
            }
            if (live) kb_v = kbn;
        };
        fill(act, 1, 1);
        __syncthreads();
        /* keys of each 12-mer's next chunk (n = fetched / 32) and the one after */
        uint32_t nk_v = (act && SWM_CHUNK < cnt_v) ? skey[lane * HK + 1] : 0u;
        uint32_t nk2_v = (act && 2 * SWM_CHUNK < cnt_v) ? skey[lane * HK + 2] : 0u;
        uint2 *cd = io.cdesc + io.bd_off[r];

        /* phase 1 (:2194-2236): the first chunk of every 12-mer, in order */
        const uint32_t get1 = act ? spc_v : 0u;
        const uint64_t has = __ballot(get1 > 0);
        const uint32_t incl = wave_incl_scan<uint32_t>(get1);
        if (get1 > 0) cd[(uint32_t)__popcll(has & below)] = make_uint2(off_v, (incl - get1) | (lane << 16));
        uint32_t nch = (uint32_t)__popcll(has), blocks = nch;
        uint32_t pos_size = rl(incl, 63);
        /* the heap (:2244-2252): 12-mers with more hits, inserted in order
         * (seeds_heap_insert :1919-1935); heap slot x lives in lane x - 1 */
        const uint32_t stage_v = (act && cnt_v > SWM_CHUNK) ? pl[off_v + SWM_CHUNK - 1].x : 0u;
        uint32_t hloc = 0, hind = 0, hsize = 0;
        for (uint64_t m = __ballot(act && spc_v != cnt_v); m; m &= m - 1ull) {
            const uint32_t i = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            hsize++;
            hloc = wl(hloc, hsize - 1, rl(stage_v, i));
            hind = wl(hind, hsize - 1, i);
            uint32_t x = hsize;
            while (x > 1) {
                const uint32_t t = x / 2;
                const uint32_t lt = rl(hloc, t - 1), lx = rl(hloc, x - 1);
                if (lt > lx) {
                    const uint32_t it = rl(hind, t - 1), ix = rl(hind, x - 1);
                    hloc = wl(wl(hloc, t - 1, lx), x - 1, lt);
                    hind = wl(wl(hind, t - 1, ix), x - 1, it);
                } else break;
                x = t;
            }
        }
        __syncthreads();   /* key buffers are read below */
        uint32_t batch = 0, cstart = 0, posoff = 0;
        /* the key read by the last pop for nk2_v of lane pend_l: from the
         * ring (pend_s, read by every pop so that it is never copied while
         * in flight) or, past the ring, from memory (pend_g) */
        uint32_t pend_s = 0, pend_g = 0, pend_l = 64u;
        bool pend_fb = false;
        /* chunk descriptors nfl .. nch - 1 of the batch wait in registers,
         * entry j in lane j % 64, and are stored 64 at a time */
        uint32_t cdx_v = 0, cdy_v = 0, nfl = nch;
        auto flush = [&]() {
            const uint32_t j = nfl + ((lane - nfl) & 63u);
            if (j < nch) cd[cstart + j] = make_uint2(cdx_v, cdy_v);
            nfl = nch;
        };
        for (;;) {
            /* the rest of this batch's chunk list (:2261-2356), wave-uniform */
            while (hsize > 0 && pos_size + SWM_CHUNK <= SWM_BATCH_POS && blocks < SWM_MAX_BLOCKS) {
                nk2_v = wl(nk2_v, pend_l, pend_fb ? pend_g : pend_s);
                const uint32_t i = rl(hind, 0);
                const uint32_t cnt = rl(cnt_v, i), pc = rl(spc_v, i), off = rl(off_v, i);
                const uint32_t c = pc / SWM_CHUNK;            /* chunk fetched now */
                const uint32_t get = min((uint32_t)SWM_CHUNK, cnt - pc);
                const uint32_t c2 = c + 2;
                pend_s = skey[i * HK + (c2 & (HK - 1))];
                cdx_v = wl(cdx_v, nch & 63u, off + pc);
                cdy_v = wl(cdy_v, nch & 63u, pos_size | (i << 16));
                nch++;
                if (nch - nfl == 64u) flush();
                pos_size += get;
                blocks++;
                spc_v = wl(spc_v, i, pc + get);
                /* seeds_heap_pop_and_insert (:1936-1962, Q6): the new root
                 * (the last entry, or this 12-mer with its next key) sifts
                 * down; its value stays in scalars and only the hole moves,
                 * which is the same sequence of swaps */
                uint32_t mloc, mind;
                if (pc + get == cnt) {
                    mloc = rl(hloc, hsize - 1);
                    mind = rl(hind, hsize - 1);
                    hsize--;
                    pend_l = 64u;
                } else {
                    mloc = rl(nk_v, i);
                    mind = i;
                    nk_v = wl(nk_v, i, rl(nk2_v, i));
                    pend_l = i;
                    pend_fb = c2 >= rl(kb_v, i) + HK;
                    if (pend_fb) pend_g = (SWM_CHUNK * c2 < cnt) ? pl[off + min(SWM_CHUNK * (c2 + 1), cnt) - 1].x : 0u;
                }
                /* while (i * 2 < size): the smaller child (the left one on a
                 * tie) moves up while it is below the new root's key; at
                 * most 5 moves in a heap of <= 64 */
                uint32_t x = 1;
#pragma unroll
                for (int lev = 0; lev < 5; lev++) {
                    const uint32_t c2x = 2 * x;
                    if (c2x >= hsize) break;
                    const uint32_t a2 = rl(hloc, c2x - 1), a3 = rl(hloc, c2x);
                    const uint32_t t = sel_gt(a2, a3, c2x + 1, c2x), lt = min(a2, a3);
                    if (lt >= mloc) break;
                    hind = wl(hind, x - 1, rl(hind, t - 1));
                    hloc = wl(hloc, x - 1, lt);
                    x = t;
                }
                hloc = wl(hloc, x - 1, mloc);
                hind = wl(hind, x - 1, mind);
            }
            nk2_v = wl(nk2_v, pend_l, pend_fb ? pend_g : pend_s);
            pend_l = 64u;
            flush();
            hq_r = wl(hq_r, nhq, r);
            hq_b = wl(hq_b, nhq, batch);
            hq_c = wl(hq_c, nhq, cstart);
            hq_n = wl(hq_n, nhq, nch | (pos_size << 8));
            hq_p = wl(hq_p, nhq, posoff);
            if (++nhq == 64u) hb_flush();
            batch++;
            cstart += nch;
            posoff += pos_size;
            if (hsize == 0) break;
            /* next batch: top up the rings if a 12-mer runs low (its keys
             * from chunk n + 2 on are read from the ring), start an empty
             * chunk list */
            {
                const uint32_t n_v = spc_v / SWM_CHUNK;
                const bool live = act && spc_v < cnt_v;
                if (__ballot(live && n_v + 2 + HK_LOW > kb_v + HK)) {
                    fill(live, n_v + 2, max(kb_v + HK, n_v + 2));
                    __syncthreads();
                }
            }
            nch = 0; nfl = 0; pos_size = 0; blocks = 0;
        }
        if (lane == 0) { io.nbatch[r] = batch; io.nuniq[r] = 0; }
        __syncthreads();
    }
    hb_flush();
}

/* The batches of heavy items, one wave each, in parallel: fetch the batch's
 * positions chunk by chunk, 15-mer filter and read-start adjustment, compact
 * in entry order at the batch's own offset; then the sorted unique list of a
 * batch of <= 64 survivors (block_heap_sort's result), larger batches to
 * k_bigsort.  The item's survivor count is the sum over its batches. */
static_assert(SWM_CHUNK == 32, "k_plan_fetch maps a chunk onto half a wave");
static_assert(SWM_BATCH_POS <= 2048, "k_bigsort sorts at most 64 x 32 values in a wave");

__global__ void __launch_bounds__(64) k_plan_fetch(const uint2 *__restrict__ pl, PlanIO io)
{
    __shared__ uint32_t sg[SWM_MAX_BLOCKS], spre[SWM_MAX_BLOCKS + 1], sseed[SWM_MAX_BLOCKS];
    __shared__ uint32_t smeta[64];
    if (*io.ovf) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t nb_all = *io.hb_count;
    /* large batches for k_bigsort wait in registers (entry j in lane j) and
     * are appended 64 at a time: one atomic on the shared counter per 64
     * batches instead of one per batch */
    uint32_t bq_r = 0, bq_b = 0, nbq = 0;
    auto big_flush = [&]() {
        uint32_t at = 0;
        if (lane == 0 && nbq) at = atomicAdd(io.big_count, nbq);
        at = (uint32_t)__builtin_amdgcn_readfirstlane((int)at);
        if (lane < nbq) io.big[at + lane] = make_uint2(bq_r, bq_b);
        nbq = 0;
    };
    for (uint32_t h = blockIdx.x; h < nb_all; h += gridDim.x) {
        const HBatch hb = io.hbatch[h];
        const uint32_t r = hb.r, bn = hb.npack & 0xffu, bpos = hb.npack >> 8;
        const uint2 *cd = io.cdesc + io.bd_off[r] + hb.cstart;
        for (uint32_t k = lane; k < bn; k += 64) {
            const uint2 d = cd[k];
            sg[k] = d.x;
            spre[k] = d.y & 0xffffu;
            sseed[k] = d.y >> 16;
        }
        if (lane == 0) spre[bn] = bpos;
        const uint32_t nsel = io.nsel[r] & ~NSEL_FLAGS;
        smeta[lane] = lane < nsel ? sel_meta(io.sel[(uint64_t)r * io.seln + lane]) : 0u;
        __syncthreads();
        uint32_t *surv = io.surv + io.surv_off[r];
        const uint32_t bstart = hb.posoff;
        uint32_t nsurv = bstart;
        uint64_t desc = 0;                  /* some chunk descends after its first entry */
        uint2 *cdw = io.cdesc + io.bd_off[r] + hb.cstart;
        /* a chunk (<= 32 entries) per half wave, lane & 31 on its entry, so
         * a load covers two chunks with consecutive addresses and lane order
         * is entry order; 16 chunks per pass, 8 loads in flight per lane;
         * kept entries are written in entry order */
        for (uint32_t k0 = 0; k0 < bn; k0 += 16) {
            uint2 v[8];
            uint32_t meta[8];
            bool live[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t k = k0 + 2u * (uint32_t)q + (lane >> 5), ei = lane & 31u;
                const bool kin = k < bn;
                const uint32_t kk = kin ? k : 0u;
                const uint32_t ent = spre[kk] + ei;
                live[q] = kin && ent < spre[kk + 1] && ent > 0;      /* entry 0: the Q7 drop */
                v[q] = pl[live[q] ? sg[kk] + ei : sg[0]];
                meta[q] = smeta[sseed[kk]];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const bool keep = live[q] && filter_pass(v[q].y, meta[q]);
                desc |= __ballot(live[q] && (lane & 31u) != 0u && (v[q].y & PL_DESCENT));
                const uint64_t m = __ballot(keep);
                if (keep) surv[nsurv + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = adjust(v[q].x, meta[q]);
                /* each chunk's survivor end (the chunk's descriptor is not
                 * read again): the blocks of k_merge */
                const uint32_t k = k0 + 2u * (uint32_t)q;
                const uint32_t lo = (uint32_t)__popcll(m & 0xffffffffull);
                if (lane == 0 && k < bn) cdw[k].x = nsurv + lo - bstart;
                if (lane == 32 && k + 1 < bn) cdw[k + 1].x = nsurv + (uint32_t)__popcll(m) - bstart;
                nsurv += (uint32_t)__popcll(m);
            }
        }
        if (desc) {
            /* block_heap_sort's merge taken literally (k_merge) */
            if (lane == 0) {
                io.bdesc[io.bd_off[r] + hb.batch] = make_uint2(bstart, nsurv - bstart);
                io.merge[atomicAdd(io.merge_count, 1u)] = h;
            }
            __syncthreads();
            continue;
        }
        /* end of batch: sorted unique list (block_heap_sort's result) */
        const uint32_t n = nsurv - bstart;
        uint2 *bd = io.bdesc + io.bd_off[r];
        uint32_t u = n;
        if (n <= 64) {
            uint32_t v = lane < n ? surv[bstart + lane] : 0xffffffffu;
            for (uint32_t kk = 2; kk <= 64; kk <<= 1) {
                for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
                    const uint32_t o = __shfl_xor(v, (int)j);
                    const bool up = (lane & kk) == 0, lower = (lane & j) == 0;
                    v = (lower == up) ? min(v, o) : max(v, o);
                }
            }
            const uint32_t prev = __shfl_up(v, 1);
            const bool head = lane < n && (lane == 0 || v != prev);
            const uint64_t m = __ballot(head);
            if (head) surv[bstart + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = v;
            u = (uint32_t)__popcll(m);
            if (lane == 0) bd[hb.batch] = make_uint2(bstart, u);
        } else {
            if (lane == 0) bd[hb.batch] = make_uint2(bstart, n);
            bq_r = wl(bq_r, nbq, r);
            bq_b = wl(bq_b, nbq, hb.batch);
            if (++nbq == 64u) big_flush();
        }
        if (lane == 0 && u) atomicAdd(&io.nuniq[r], u);
        __syncthreads();
    }
    big_flush();
}

/* Heavy batches that fetched a descending chunk (listed by k_plan_fetch,
 * which left the batch's survivors in fetch order and each chunk's survivor
 * end in its descriptor): block_heap_sort's merge taken literally, one
 * workgroup per batch, the survivors staged in LDS and merged by one lane. */
__global__ void __launch_bounds__(64) k_merge(PlanIO io)
{
    __shared__ uint32_t s_pos[SWM_BATCH_POS];
    __shared__ uint32_t s_ind[2 * SWM_MAX_BLOCKS + 4];
    __shared__ uint32_t s_hp[SWM_MAX_BLOCKS + 2], s_hb[SWM_MAX_BLOCKS + 2];
    if (*io.ovf) return;
    const uint32_t nm = *io.merge_count;
    for (uint32_t i = blockIdx.x; i < nm; i += gridDim.x) {
        const HBatch hb = io.hbatch[io.merge[i]];
        const uint32_t r = hb.r, bn = hb.npack & 0xffu;
        const uint2 *cd = io.cdesc + io.bd_off[r] + hb.cstart;
        uint2 *bd = io.bdesc + io.bd_off[r] + hb.batch;
        const uint2 b = *bd;
        uint32_t *surv = io.surv + io.surv_off[r] + b.x;
        for (uint32_t j = threadIdx.x; j < b.y; j += 64) s_pos[j] = surv[j];
        for (uint32_t k = threadIdx.x; k < bn; k += 64) {
            s_ind[2 * k] = k ? cd[k - 1].x : 0u;
            s_ind[2 * k + 1] = cd[k].x;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t u = block_merge(s_pos, s_ind, bn, s_hp, s_hb, surv);
            *bd = make_uint2(b.x, u);
            if (u) atomicAdd(&io.nuniq[r], u);
        }
        __syncthreads();
    }
}

/* One wave per large batch (65 .. 2048 values), sorted in registers: R =
 * 2 .. 32 values per lane (the batch padded to 64 R with ~0), logical
 * element i = lane * R + r.  A bitonic network in its "flip" form -- the
 * first comparator of each merge phase pairs i with its mirror in the
 * block, the rest pair i with i ^ j -- keeps every comparator ascending,
 * so no direction logic: a comparator inside a lane is a min and a max, one
 * across lanes is a lane permute and one v_med3 with 0 (keep the smaller)
 * or ~0 (keep the larger) by the lane's side.  Lane permutes are DPP for
 * the masks within a row of 16 (quad_perm, row_mirror, row_half_mirror,
 * row_ror:8), ds_swizzle for 16 and 31, ds_bpermute only for 32 and 63.
 * Then unique: element i is kept when i < n and it differs from i - 1. */
template <int M>
__device__ __forceinline__ uint32_t xlane(uint32_t x)
{
    if constexpr (M == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xf, 0xf, false);  /* quad_perm 1,0,3,2 */
    else if constexpr (M == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xf, 0xf, false);
    else if constexpr (M == 3) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x1B, 0xf, 0xf, false);
    else if constexpr (M == 4)                                                              /* ^3 then ^7 */
        return (uint32_t)__builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp((int)x, 0x1B, 0xf, 0xf, false),
                                                  0x141, 0xf, 0xf, false);
    else if constexpr (M == 7) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x141, 0xf, 0xf, false); /* row_half_mirror */
    else if constexpr (M == 8) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xf, 0xf, false); /* row_ror:8 */
    else if constexpr (M == 15) return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x140, 0xf, 0xf, false); /* row_mirror */
    else if constexpr (M == 16) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x401F);  /* xor 16 within 32 */
    else if constexpr (M == 31) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)x, 0x7C1F);  /* xor 31 within 32 */
    else return (uint32_t)__shfl_xor((int)x, M, 64);
}

__device__ __forceinline__ uint32_t med3u(uint32_t a, uint32_t b, uint32_t c)
{
    return min(max(a, b), max(min(a, b), c));      /* v_med3_u32 */
}

/* half-cleaner steps: i with i ^ 2^B, B .. 0 */
template <int R, int RB, int B>
__device__ __forceinline__ void bitonic_steps(uint32_t (&v)[R], uint32_t lane)
{
    if constexpr (B >= 0) {
        if constexpr (B < RB) {
            constexpr int j = 1 << B;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if ((r & j) == 0) {
                    const uint32_t lo = min(v[r], v[r | j]), hi = max(v[r], v[r | j]);
                    v[r] = lo;
                    v[r | j] = hi;
                }
            }
        } else {
            const uint32_t c = ((lane >> (B - RB)) & 1u) ? 0xffffffffu : 0u;
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = med3u(v[r], xlane<(1 << (B - RB))>(v[r]), c);
        }
        bitonic_steps<R, RB, B - 1>(v, lane);
    }
}

/* merge phases K .. RB + 6: the flip step (i with its mirror in a block of
 * 2^K), then the half-cleaners */
template <int R, int RB, int K>
__device__ __forceinline__ void bitonic_phases(uint32_t (&v)[R], uint32_t lane)
{
    if constexpr (K <= RB + 6) {
        if constexpr (K <= RB) {
            constexpr int mk = (1 << K) - 1;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int q = r ^ mk;
                if (r < q) {
                    const uint32_t lo = min(v[r], v[q]), hi = max(v[r], v[q]);
                    v[r] = lo;
                    v[q] = hi;
                }
            }
        } else {
            const uint32_t c = ((lane >> (K - RB - 1)) & 1u) ? 0xffffffffu : 0u;
#pragma unroll
            for (int r = 0; r < R / 2; r++) {
                const uint32_t a = xlane<(1 << (K - RB)) - 1>(v[R - 1 - r]);
                const uint32_t b = xlane<(1 << (K - RB)) - 1>(v[r]);
                v[r] = med3u(v[r], a, c);
                v[R - 1 - r] = med3u(v[R - 1 - r], b, c);
            }
        }
        bitonic_steps<R, RB, K - 2>(v, lane);
        bitonic_phases<R, RB, K + 1>(v, lane);
    }
}

template <int R>
__device__ __forceinline__ void wave_bitonic(uint32_t (&v)[R], uint32_t lane)
{
    constexpr int RB = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : R == 16 ? 4 : 5;
    bitonic_phases<R, RB, 1>(v, lane);
}

/* sort base[0 .. n) (n <= 64 R) and write its distinct values in place;
 * returns their count */
template <int R>
__device__ __forceinline__ uint32_t wave_sort_unique(uint32_t *base, uint32_t n, uint32_t lane)
{
    uint32_t v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = (uint32_t)r * 64u + lane;
        v[r] = i < n ? base[i] : 0xffffffffu;
    }
    wave_bitonic<R>(v, lane);
    /* element i - 1 of element lane * R: the last of the lane below (wave_shr:1) */
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[R - 1], 0x138, 0xf, 0xf, false);
    uint32_t cnt = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = lane * R + (uint32_t)r;
        const uint32_t left = r ? v[r - 1] : prev;
        cnt += (i < n && (i == 0 || v[r] != left)) ? 1u : 0u;
    }
    const uint32_t incl = wave_incl_scan_dpp(cnt);
    uint32_t wp = incl - cnt;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const uint32_t i = lane * R + (uint32_t)r;
        const uint32_t left = r ? v[r - 1] : prev;
        if (i < n && (i == 0 || v[r] != left)) base[wp++] = v[r];
    }
    return (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
}

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) k_bigsort(const uint2 *__restrict__ big, const uint32_t *__restrict__ big_count,
                                                const uint64_t *__restrict__ surv_off,
                                                const uint32_t *__restrict__ bd_off, uint32_t *__restrict__ surv,
                                                uint2 *__restrict__ bdesc, uint32_t *__restrict__ nuniq)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t nb = *big_count;
    for (uint32_t bt = blockIdx.x; bt < nb; bt += gridDim.x) {
        const uint2 rb = big[bt];
        uint2 *dp = bdesc + bd_off[rb.x] + rb.y;
        const uint2 d = *dp;
        const uint32_t n = d.y;
        uint32_t *base = surv + surv_off[rb.x] + d.x;
        uint32_t total;
        if (n <= 128) total = wave_sort_unique<2>(base, n, lane);
        else if (n <= 256) total = wave_sort_unique<4>(base, n, lane);
        else if (n <= 512) total = wave_sort_unique<8>(base, n, lane);
        else if (n <= 1024) total = wave_sort_unique<16>(base, n, lane);
        else total = wave_sort_unique<32>(base, n, lane);
        if (lane == 0) {
            dp->y = total;
            atomicSub(&nuniq[rb.x], n - total);
        }
    }
}

/* --------------------------------------------------------------- k_compact */

struct CandSet {
    uint32_t *val = nullptr;      /* read-start candidate */
    uint32_t *rs = nullptr;       /* (read << 1) | strand */
    uint32_t *bi = nullptr;       /* batch << 12 | index within batch */
    int32_t *err = nullptr;       /* -2: not verified (pos_start rejected) */
    int32_t *loc = nullptr;
    uint32_t *rec = nullptr;      /* record location (cigar->location) for passing lanes */
    uint32_t *info = nullptr;     /* CIGAR bytes | kind bit 16 (1 = traceback) | bit 17: record location < 0 */
    const uint32_t *ovf = nullptr;    /* capacity overflow flag (see PlanIO) */
    /* record output (null unless records are asked for): per candidate,
     * WR_REC if the emission writes its record (plus WR_BLOCK when that
     * record opens one of put_cigar's output blocks), the bytes of its CIGAR
     * text, and the traceback's merged runs */
    uint64_t *wr = nullptr;
    uint32_t *tb = nullptr;
    uint16_t *runs = nullptr;     /* [ci * RUN_SLOT]: count, then runs in traceback order */
    /* verification order (null: candidate order): perm[p] is the candidate
     * the Myers kernels take at place p, unless *perm_id is set (one read
     * length in the batch: the identity).  perm_id is 8-byte aligned with
     * the overflow flag right after it (perm_id + 1 == ovf), so the Myers
     * kernels load both in one 64-bit load */
    const uint32_t *perm = nullptr;
    const uint32_t *perm_id = nullptr;
};

/* A traceback's CIGAR runs, M and S runs merged as the emission merges them
 * (mapper_cpe.c:479-507): at most 2e+1 <= 31 entries (an I/D run or an S
 * costs an error, the M groups sit between them), each size | kind << 8. */
#define RUN_SLOT 32
#define RUN_M 0u
#define RUN_I 1u
#define RUN_D 2u

#define INFO_TRACE (1u << 16)
#define INFO_NEG (1u << 17)

/* The records path's per-candidate word: one scan over it gives a written
 * record's rank (low half) and its output block's rank (high half). */
#define WR_REC 1ull
#define WR_BLOCK (1ull << 32)

/* (identity order, overflow) in one load: a workgroup of the Myers grid,
 * which covers every candidate, starts on one round trip, not two.  The
 * caller stands down on myers_stop(): the identity flag is 0 or 1, so its
 * bit 31 never counts, but testing it keeps the compiler from deferring
 * that word's load past the branch. */
__device__ __forceinline__ uint2 myers_flags(const CandSet &c)
{
    return *reinterpret_cast<const uint2 *>(c.perm_id);
}
__device__ __forceinline__ bool myers_stop(uint2 fl) { return ((fl.x >> 31) | fl.y) != 0u; }

/* Verification order by read length.  In candidate order a lane's pair
 * straddles two reads of different lengths wherever a read's candidate
 * count is odd (the packed kernel then runs both halves alone), and a wave
 * holds reads of many lengths whose steps all wait for the longest.  Three
 * small kernels lay the candidate indices out by read length instead: per
 * length a bin, per (read, strand) one run of its candidates.  A batch of
 * one length keeps the identity (the scatter stands down).  The order
 * changes no result: each candidate writes only its own slots. */
#define LSORT_WG_PER_CU 4

__device__ __forceinline__ uint32_t lbin_of(uint32_t L) { return L < 255u ? L : 255u; }

/* n candidates to LDS bin b, per wave: one atomic when the wave's items
 * (those with candidates) share one length, else one per lane (same-address
 * LDS atomics serialise: one length per batch made a lane-per-atomic count
 * take 19 us of a 1M-read step).  Every lane of the wave calls it. */
__device__ __forceinline__ void lhist_add(uint32_t *h, uint32_t n, uint32_t b)
{
    const uint64_t has = __ballot(n != 0u);
    if (!has) return;
    const uint32_t b0 = (uint32_t)__shfl((int)b, __builtin_ctzll(has), 64);
    if (__all(n == 0u || b == b0)) {
        uint32_t t = n;
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) t += (uint32_t)__shfl_xor((int)t, d, 64);
        if (__lane_id() == 0) atomicAdd(&h[b0], t);
    } else if (n) {
        atomicAdd(&h[b], n);
    }
}

/* The workgroup's items (grid-stride, stride = the grid's threads) into its
 * LDS length bins, the fields of LC_PER items per thread loaded together
 * (one item at a time left the count latency-bound: 21 us of a 1M-read
 * step). */
#define LC_PER 8
__device__ __forceinline__ void lhist_count(uint32_t *h, uint32_t nrs, const uint32_t *__restrict__ cand_off,
                                            const uint32_t *__restrict__ lens, uint64_t stride)
{
    for (uint64_t r0 = (uint64_t)blockIdx.x * blockDim.x; r0 < nrs; r0 += LC_PER * stride) {
        /* unconditional loads (a place past the end reads item 0 and counts
         * nothing), so that no load waits on a branch */
        uint32_t aq[LC_PER], cq[LC_PER], lq[LC_PER];
#pragma unroll
        for (int k = 0; k < LC_PER; k++) {
            const uint64_t r = r0 + (uint64_t)k * stride + threadIdx.x;
            const uint64_t rr = r < nrs ? r : 0u;
            aq[k] = cand_off[rr];
            cq[k] = cand_off[rr + 1];
            lq[k] = lens[rr >> 1];
        }
#pragma unroll
        for (int k = 0; k < LC_PER; k++) {
            const bool in = r0 + (uint64_t)k * stride + threadIdx.x < nrs;
            lhist_add(h, in ? cq[k] - aq[k] : 0u, lbin_of(lq[k]));
        }
    }
}

/* candidates per read length, one atomic per length per workgroup */
__global__ void __launch_bounds__(256) k_lcount(uint32_t nrs, const uint32_t *__restrict__ cand_off,
                                                const uint32_t *__restrict__ lens, const uint32_t *__restrict__ ovf,
                                                uint32_t *__restrict__ bins)
{
    __shared__ uint32_t h[256];
    if (*ovf) return;
    const uint32_t tid = threadIdx.x;
    h[tid] = 0;
    __syncthreads();
    lhist_count(h, nrs, cand_off, lens, (uint64_t)gridDim.x * 256u);
    __syncthreads();
    if (h[tid]) atomicAdd(&bins[(blockIdx.x % LBIN_ROWS) * 256u + tid], h[tid]);
}

/* one workgroup: each length's first place (bins[LBIN_ROWS * 256 + L], the
 * rows summed), and the identity flag when at most one length is used */
__global__ void __launch_bounds__(256) k_lscan(const uint32_t *__restrict__ ovf, uint32_t *__restrict__ bins,
                                               uint32_t *__restrict__ ident)
{
    __shared__ uint32_t s_w[4];
    if (*ovf) return;
    const uint32_t tid = threadIdx.x;
    uint32_t v = 0;
#pragma unroll
    for (uint32_t w = 0; w < LBIN_ROWS; w++) v += bins[w * 256u + tid];
    uint32_t tot, used;
    const uint32_t ex = block_incl_scan<uint32_t, 256>(v, s_w, &tot) - v;
    (void)block_incl_scan<uint32_t, 256>(v ? 1u : 0u, s_w, &used);
    bins[LBIN_ROWS * 256u + tid] = ex;
    if (tid == 0) *ident = used <= 1u ? 1u : 0u;
}

/* the same items per workgroup as k_lcount: its share of each bin (one
 * atomic per length), then every item's run at its place in the share;
 * runs of up to 16 written by their own lane, longer ones by the wave */
__global__ void __launch_bounds__(256) k_lscatter(uint32_t nrs, const uint32_t *__restrict__ cand_off,
                                                  const uint32_t *__restrict__ lens, const uint32_t *__restrict__ ovf,
                                                  const uint32_t *__restrict__ ident, uint32_t *__restrict__ bins,
                                                  uint32_t *__restrict__ perm)
{
    __shared__ uint32_t h[256], base[256];
    if (*ovf || *ident) return;
    const uint32_t tid = threadIdx.x, lane = __lane_id();
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    h[tid] = 0;
    __syncthreads();
    lhist_count(h, nrs, cand_off, lens, stride);
    __syncthreads();
    base[tid] = h[tid] ? atomicAdd(&bins[LBIN_ROWS * 256u + tid], h[tid]) : 0u;
    h[tid] = 0;
    __syncthreads();
    for (uint64_t r00 = (uint64_t)blockIdx.x * 256u; r00 < nrs; r00 += LC_PER * stride) {
        /* the fields of LC_PER items loaded together, as in the count */
        uint32_t aq[LC_PER], cq[LC_PER], lq[LC_PER];
#pragma unroll
        for (int q = 0; q < LC_PER; q++) {
            const uint64_t r = r00 + (uint64_t)q * stride + tid;
            const uint64_t rr = r < nrs ? r : 0u;
            aq[q] = cand_off[rr];
            cq[q] = cand_off[rr + 1];
            lq[q] = lens[rr >> 1];
        }
#pragma unroll
        for (int q = 0; q < LC_PER; q++) {
            const bool in = r00 + (uint64_t)q * stride + tid < nrs;
            const uint32_t start = aq[q], n = in ? cq[q] - aq[q] : 0u;
            uint32_t pos = 0;
            if (n) {
                const uint32_t b = lbin_of(lq[q]);
                pos = base[b] + atomicAdd(&h[b], n);
            }
            const bool own = n <= 16u;
            if (own)
                for (uint32_t j = 0; j < n; j++) perm[pos + j] = start + j;
            uint64_t m = __ballot(!own);
            while (m) {
                const int k = __builtin_ctzll(m);
                m &= m - 1ull;
                const uint32_t pk = (uint32_t)__shfl((int)pos, k, 64), sk = (uint32_t)__shfl((int)start, k, 64);
                const uint32_t nk = (uint32_t)__shfl((int)n, k, 64);
                for (uint32_t j = lane; j < nk; j += 64u) perm[pk + j] = sk + j;
            }
        }
    }
}

/* Items with more than heavy_cands candidates go to the cooperative kernels. */
/* The other items, 256 per workgroup, copied flat: their batch descriptors
 * laid end to end (a scan of the batch counts) and loaded together, the
 * survivor counts scanned in LDS, then every survivor copied by the thread
 * of its place in the workgroup's span.  (A thread per item walking its
 * batches four survivors per memory round trip left each wave waiting on
 * its longest item.)  Items with more than heavy_cands candidates go to the
 * heavy list for k_compact_heavy. */
#define CPT_SLOTS 512   /* batch descriptors staged at a time */
__global__ void __launch_bounds__(256) k_compact(uint32_t nrs, const uint64_t *__restrict__ surv_off,
                                                 const uint32_t *__restrict__ surv, const uint32_t *__restrict__ bd_off,
                                                 const uint2 *__restrict__ bdesc, const uint32_t *__restrict__ nbatch,
                                                 const uint32_t *__restrict__ cand_off, CandSet c, uint32_t heavy_cands,
                                                 uint32_t *__restrict__ heavy, uint32_t *__restrict__ heavy_count)
{
    __shared__ uint32_t s_tmp[2], s_w[4];
    __shared__ uint32_t s_bpre[257], s_bdo[256], s_o[256];   /* per item: first slot, descriptors, next output */
    __shared__ uint32_t s_cn[256];                            /* per item: its candidate count */
    __shared__ uint64_t s_so[256];                            /* per item: first survivor */
    __shared__ uint32_t s_it[CPT_SLOTS], s_pre[CPT_SLOTS + 1], s_ob[CPT_SLOTS], s_bj[CPT_SLOTS];
    __shared__ uint64_t s_src[CPT_SLOTS];
    const uint32_t tid = threadIdx.x;
    const uint32_t R0 = blockIdx.x * 256u;
    const uint32_t r = R0 + tid;
    const bool live = r < nrs;
    /* the item's fields are loaded with the overflow flag (they lie in
     * buffers sized by the item count, whatever the capacities): one
     * memory round trip, then one for the descriptors and one for the
     * survivors */
    const uint32_t ovf = *c.ovf;
    uint32_t co = 0, cn = 0, nbr = 0, bdo = 0;
    uint64_t so = 0;
    if (live) {
        co = cand_off[r];
        cn = cand_off[r + 1] - co;
        nbr = nbatch[r];
        bdo = bd_off[r];
        so = surv_off[r];
    }
    if (ovf) return;
    const bool big = cn > heavy_cands;
    block_append(big, r, heavy, heavy_count, s_tmp);
    /* an item without candidates has only empty batches: nothing to copy */
    const uint32_t nb = (live && !big && cn) ? nbr : 0u;
    uint32_t tb;
    const uint32_t bincl = block_incl_scan<uint32_t, 256>(nb, s_w, &tb);
    s_bpre[tid] = bincl - nb;
    if (tid == 0) s_bpre[256] = tb;
    s_o[tid] = co;
    s_cn[tid] = cn;
    s_bdo[tid] = bdo;
    s_so[tid] = so;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < tb; c0 += CPT_SLOTS) {
        const uint32_t n = min((uint32_t)CPT_SLOTS, tb - c0);
        /* slots 2 tid and 2 tid + 1: the owner is the last item whose first
         * slot is not past it (items without batches share their first slot
         * with the next item and come before it) */
        uint32_t cq[2] = {0u, 0u};
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t k = 2 * tid + (uint32_t)q;
            if (k < n) {
                const uint32_t slot = c0 + k;
                uint32_t lo = 0, hi = 255;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (s_bpre[mid] <= slot) lo = mid;
                    else hi = mid - 1;
                }
                const uint32_t b = slot - s_bpre[lo];
                /* an item's only batch starts its survivors and holds all
                 * its candidates: no descriptor load */
                const uint2 d = s_bpre[lo + 1] - s_bpre[lo] == 1u ? make_uint2(0u, s_cn[lo]) : bdesc[s_bdo[lo] + b];
                s_it[k] = lo;
                s_src[k] = s_so[lo] + d.x;
                s_bj[k] = b << 12;
                cq[q] = d.y;
            }
        }
        const uint32_t pair = cq[0] + cq[1];
        uint32_t te;
        const uint32_t pex = block_incl_scan<uint32_t, 256>(pair, s_w, &te) - pair;
        if (2 * tid < n) s_pre[2 * tid] = pex;
        if (2 * tid + 1 < n) s_pre[2 * tid + 1] = pex + cq[0];
        if (tid == 0) s_pre[n] = te;
        __syncthreads();
        /* output index of each slot's first survivor: its item's next output
         * plus the survivors of the item's earlier slots in this chunk */
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const uint32_t k = 2 * tid + (uint32_t)q;
            if (k < n) {
                const uint32_t it = s_it[k];
                const uint32_t first = s_bpre[it] > c0 ? s_bpre[it] - c0 : 0u;
                s_ob[k] = s_o[it] + s_pre[k] - s_pre[first];
            }
        }
        __syncthreads();
        if (nb) {
            const uint32_t f = s_bpre[tid], l = f + nb;
            const uint32_t a = min(max(f, c0), c0 + n) - c0, z = min(max(l, c0), c0 + n) - c0;
            s_o[tid] += s_pre[z] - s_pre[a];
        }
        /* survivor e of the chunk: slot = the last one whose prefix is not
         * past e (empty slots share their prefix with the next), four
         * loads in flight per thread */
        for (uint32_t e0 = tid; e0 < te; e0 += 1024) {
            uint32_t kq[4], jq[4], vq[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t e = e0 + 256u * (uint32_t)q;
                kq[q] = 0;
                jq[q] = 0;
                vq[q] = 0;
                if (e < te) {
                    uint32_t lo = 0, hi = n - 1;
                    while (lo < hi) {
                        const uint32_t mid = (lo + hi + 1) >> 1;
                        if (s_pre[mid] <= e) lo = mid;
                        else hi = mid - 1;
                    }
                    kq[q] = lo;
                    jq[q] = e - s_pre[lo];
                    vq[q] = surv[s_src[lo] + jq[q]];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t e = e0 + 256u * (uint32_t)q;
                if (e < te) {
                    const uint32_t out = s_ob[kq[q]] + jq[q];
                    c.val[out] = vq[q];
                    c.rs[out] = R0 + s_it[kq[q]];
                    c.bi[out] = s_bj[kq[q]] | jq[q];
                }
            }
        }
        __syncthreads();
    }
}

/* One workgroup per (read, strand) with many candidates. */
__global__ void __launch_bounds__(256) k_compact_heavy(const uint32_t *__restrict__ heavy,
                                                       const uint32_t *__restrict__ heavy_count,
                                                       const uint64_t *__restrict__ surv_off,
                                                       const uint32_t *__restrict__ surv,
                                                       const uint32_t *__restrict__ bd_off,
                                                       const uint2 *__restrict__ bdesc,
                                                       const uint32_t *__restrict__ nbatch,
                                                       const uint32_t *__restrict__ cand_off, CandSet c)
{
    /* up to 256 batches at a time: their descriptors loaded at once and
     * their counts scanned in LDS, then the span copied flat (thread t on
     * elements t, t + 256, ..., four loads in flight, its batch found by a
     * pointer that only moves forward), so no batch waits on its own
     * descriptor load */
    __shared__ uint32_t s_x[256], s_pre[257], s_w[4];
    const uint32_t tid = threadIdx.x;
    const uint32_t nh = *heavy_count;
    for (uint32_t h = blockIdx.x; h < nh; h += gridDim.x) {
        const uint32_t r = heavy[h];
        const uint32_t *s = surv + surv_off[r];
        const uint2 *bd = bdesc + bd_off[r];
        uint32_t o = cand_off[r];
        const uint32_t nb = nbatch[r];
        for (uint32_t b0 = 0; b0 < nb; b0 += 256) {
            const uint32_t nbc = min(256u, nb - b0);
            const uint2 d = tid < nbc ? bd[b0 + tid] : make_uint2(0u, 0u);
            uint32_t total;
            const uint32_t incl = block_incl_scan<uint32_t, 256>(d.y, s_w, &total);
            s_x[tid] = d.x;
            s_pre[tid] = incl - d.y;
            if (tid == 0) s_pre[nbc] = total;
            __syncthreads();
            uint32_t k = 0;
            for (uint32_t e0 = tid; e0 < total; e0 += 1024) {
                /* four elements per thread, their loads issued together */
                uint32_t kq[4], jq[4], vq[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t e = e0 + 256u * (uint32_t)q;
                    if (e < total) {
                        while (s_pre[k + 1] <= e) k++;
                        kq[q] = k;
                        jq[q] = e - s_pre[k];
                        vq[q] = s[s_x[k] + jq[q]];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t e = e0 + 256u * (uint32_t)q;
                    if (e < total) {
                        c.val[o + e] = vq[q];
                        c.rs[o + e] = r;
                        c.bi[o + e] = ((b0 + kq[q]) << 12) | jq[q];
                    }
                }
            }
            o += total;
            __syncthreads();
        }
    }
}

/* ----------------------------------------------------------------- k_myers */

/* Banded Myers for one candidate (the recurrence of mapper_cpe.c:570-631 with
 * the end scan of :634-650).  The window planes are aligned at the window
 * start; at read base t the match mask over band bits j = window base t + j
 * is built from the plane bits (the compiled kernel's form, :845-857).  With
 * STORE the per-step D0 and HP words are kept for the traceback. */
#define SWM_NW 9   /* window words: ((254 - 1) >> 5) + 2 */

template <bool STORE>
__device__ __forceinline__ void myers_lane(const uint32_t *__restrict__ rw, uint32_t L, uint32_t e,
                                           uint32_t ref_len, uint32_t (&wl)[SWM_NW], uint32_t (&wh)[SWM_NW],
                                           int32_t *err_out, int32_t *loc_out, uint32_t *d0a, uint32_t *hpa)
{
    /* the read's 2-bit words are loaded once and rotate with the window, so
     * the step loop touches no memory */
    uint32_t rv[16];
#pragma unroll
    for (int j = 0; j < 16; j++) rv[j] = rw[j];
    const uint32_t B = 2 * e + 1;
    const uint32_t mask = (1u << B) - 1u;
    uint32_t vp = 0, vn = 0, ones = 0;
    const uint32_t nfull = L >> 5, rem = L & 31u;

#define SWM_STEP(S, R, T)                                                                  \
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
        if (STORE) { d0a[(T)] = d0; hpa[(T)] = hp; }                                       \
    }

    for (uint32_t c = 0; c < nfull; c++) {
        const uint32_t r0 = rv[0], r1 = rv[1];
        const uint32_t l0 = wl[0], l1 = wl[1], h0 = wh[0], h1 = wh[1];
#pragma unroll
        for (int s = 0; s < 16; s++) SWM_STEP(s, r0, 32 * c + s)
#pragma unroll
        for (int s = 16; s < 32; s++) SWM_STEP(s, r1, 32 * c + s)
#pragma unroll
        for (int j = 0; j < SWM_NW - 1; j++) { wl[j] = wl[j + 1]; wh[j] = wh[j + 1]; }
#pragma unroll
        for (int j = 0; j < 14; j++) rv[j] = rv[j + 2];
    }
    if (rem) {
        const uint32_t r0 = rv[0], r1 = rv[1];
        const uint32_t l0 = wl[0], l1 = wl[1], h0 = wh[0], h1 = wh[1];
#pragma unroll
        for (int s = 0; s < 16; s++) {
            if ((uint32_t)s < rem) SWM_STEP(s, r0, 32 * nfull + s)
        }
#pragma unroll
        for (int s = 16; s < 32; s++) {
            if ((uint32_t)s < rem) SWM_STEP(s, r1, 32 * nfull + s)
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
    /* all SWM_NW + 1 pairs, unconditionally: a window starts at a genome
     * position, and the device genome has SWM_WIN_PAIRS_PAD zero pairs past
     * its end, so the loads stay in bounds and go out together */
    const uint2 *P = reinterpret_cast<const uint2 *>(b2) + (win_start >> 5);
    const uint32_t sh = win_start & 31u;
    (void)L;
    uint2 pr[SWM_NW + 1];
#pragma unroll
    for (int j = 0; j < SWM_NW + 1; j++) pr[j] = P[j];
#pragma unroll
    for (int j = 0; j < SWM_NW; j++) {
        wl[j] = __builtin_amdgcn_alignbit(pr[j + 1].x, pr[j].x, sh);
        wh[j] = __builtin_amdgcn_alignbit(pr[j + 1].y, pr[j].y, sh);
    }
}

/* The candidate count is read on the device (the last entry of the candidate
 * offsets), so a fixed grid serves any batch: grid-stride loop. */
/* Per workgroup, what verification hands on beside each candidate's
 * distance: the candidates that need the traceback (appended to the global
 * list once per round of the workgroup: one global atomic, the entries
 * placed through LDS), and the band steps (read length per verified
 * window) and rejected windows, summed and added once at the end (same-
 * address atomics from every wave would serialise at one L2 channel).
 * Every thread of the workgroup calls round() and finish(). */
struct VerifyTally {
    uint32_t *list, *count;
    unsigned long long *n_steps;
    uint32_t *n_rej;
    uint32_t *s_list, *s_n, *s_base, *s_rej;
    unsigned long long *s_steps;
    unsigned long long steps;
    uint32_t rej;
    __device__ void start()
    {
        steps = 0;
        rej = 0;
        if (threadIdx.x == 0) { *s_n = 0; *s_steps = 0; *s_rej = 0; }
        __syncthreads();
    }
    /* a and b: this thread's traceback candidates of the round (~0u: none) */
    __device__ void round(uint32_t a, uint32_t b)
    {
        const uint32_t lane = __lane_id();
        const uint64_t below = (1ull << lane) - 1ull;
        const uint64_t ma = __ballot(a != ~0u), mb = __ballot(b != ~0u);
        uint32_t wb = 0;
        if (lane == 0 && (ma | mb)) wb = atomicAdd(s_n, (uint32_t)(__popcll(ma) + __popcll(mb)));
        wb = (uint32_t)__shfl((int)wb, 0, 64);
        if (a != ~0u) s_list[wb + (uint32_t)__popcll(ma & below)] = a;
        if (b != ~0u) s_list[wb + (uint32_t)__popcll(ma) + (uint32_t)__popcll(mb & below)] = b;
        __syncthreads();
        const uint32_t n = *s_n;
        if (n) {
            if (threadIdx.x == 0) *s_base = atomicAdd(count, n);
            __syncthreads();
            const uint32_t base = *s_base;
            for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) list[base + k] = s_list[k];
            __syncthreads();
            if (threadIdx.x == 0) *s_n = 0;
            __syncthreads();
        }
    }
    __device__ void finish()
    {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)steps, d, 64);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(steps >> 32), d, 64);
            steps += ((unsigned long long)hi << 32) | lo;
            rej += (uint32_t)__shfl_xor((int)rej, d, 64);
        }
        if (__lane_id() == 0) {
            if (steps) atomicAdd(s_steps, steps);
            if (rej) atomicAdd(s_rej, rej);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            if (*s_steps) atomicAdd(n_steps, *s_steps);
            if (*s_rej) atomicAdd(n_rej, *s_rej);
        }
    }
};
#define VT_LIST 512    /* traceback entries per round: two per thread of a workgroup of at most 256 */

/* Hamming check of align_reads (:252-287) for a verified candidate with
 * err <= e: the read against the window at the implied start loc - L + 1.
 * Equal counts give a "<L>M" record; the rest go to the traceback (returns
 * true).  The window is loaded again (the verification rotated its words
 * out of the registers; its lines are still in the cache). */
__device__ __forceinline__ bool ham_window(uint32_t i, int32_t er, int32_t lc, uint32_t st, uint32_t L,
                                           const uint32_t *__restrict__ w, const uint32_t *__restrict__ b2,
                                           const CandSet &c)
{
    uint32_t wl[SWM_NW], wh[SWM_NW];
    load_window(b2, st, L, wl, wh);
    const int sl = lc - (int)L + 1;
    /* branch-free over all words (a mask of 0 past the read), so every
     * read plane load is issued up front */
    int tmp = 0;
#pragma unroll
    for (int ii = 0; ii < SWM_NW - 1; ii++) {
        const uint32_t lo = 32u * (uint32_t)ii;
        const uint32_t left = L > lo ? L - lo : 0u;
        const uint32_t mask = left >= 32u ? 0xffffffffu : ((1u << left) - 1u);
        const uint32_t lref = __builtin_amdgcn_alignbit(wl[ii + 1], wl[ii], (uint32_t)sl);
        const uint32_t href = __builtin_amdgcn_alignbit(wh[ii + 1], wh[ii], (uint32_t)sl);
        tmp += __popc(((w[17 + 2 * ii] ^ href) | (w[16 + 2 * ii] ^ lref)) & mask);
    }
    if (tmp == er) {
        c.rec[i] = st + (uint32_t)sl + 1u;
        c.info[i] = 2u + (uint32_t)ndigits(L) + 1u;
        return false;
    }
    c.rec[i] = 0;
    c.info[i] = INFO_TRACE;
    return true;
}

template <bool PERM>
__device__ __forceinline__ void myers_loop(uint32_t ncand, const uint32_t *__restrict__ perm,
                                           const uint32_t *__restrict__ lens, const uint32_t *__restrict__ words,
                                           const uint32_t *__restrict__ b2, uint64_t ref_length, uint32_t e,
                                           const CandSet &c, VerifyTally &vt)
{
    for (uint32_t p0 = blockIdx.x * blockDim.x; p0 < ncand; p0 += gridDim.x * blockDim.x) {
        const uint32_t p = p0 + threadIdx.x;
        uint32_t tq = ~0u;
        if (p < ncand) {
            const uint32_t i = PERM ? perm[p] : p;    /* place p of the verification order */
            const uint32_t rs = c.rs[i];
            const uint32_t L = lens[rs >> 1];
            uint32_t ref_len;
            const uint32_t st = pos_start(ref_length, e, c.val[i], L, &ref_len);
            if (st == 0xffffffffu) {
                c.err[i] = -2;
                c.loc[i] = -1;
                vt.rej++;                                /* pos_start rejected the window: not a lane */
            } else {
                uint32_t wl[SWM_NW], wh[SWM_NW];
                load_window(b2, st, L, wl, wh);
                const uint32_t *w = words + (uint64_t)rs * 32;
                int32_t er, lc;
                myers_lane<false>(w, L, e, ref_len, wl, wh, &er, &lc, nullptr, nullptr);
                c.err[i] = er;
                c.loc[i] = lc;
                vt.steps += L;
                if (er <= (int32_t)e && ham_window(i, er, lc, st, L, w, b2, c)) tq = i;
            }
        }
        vt.round(tq, ~0u);
    }
}

/* The verification kernels also run the Hamming check of every passing
 * candidate (its window is in registers), list the traceback candidates and
 * count band steps and rejected windows (the run's statistics). */
#define VT_ARGS uint32_t *__restrict__ trace, uint32_t *__restrict__ trace_count, \
                unsigned long long *__restrict__ n_steps, uint32_t *__restrict__ n_rej
#define VT_SETUP                                                                           \
    __shared__ uint32_t s_list[VT_LIST], s_n, s_base, s_rej;                               \
    __shared__ unsigned long long s_steps;                                                 \
    VerifyTally vt{trace, trace_count, n_steps, n_rej, s_list, &s_n, &s_base, &s_rej, &s_steps, 0, 0}; \
    vt.start();

__global__ void __launch_bounds__(256) k_myers(const uint32_t *__restrict__ ncand_p, const uint32_t *__restrict__ lens,
                                               const uint32_t *__restrict__ words, const uint32_t *__restrict__ b2,
                                               uint64_t ref_length, uint32_t e, CandSet c, VT_ARGS)
{
    const uint2 fl = myers_flags(c);
    if (myers_stop(fl)) return;
    const uint32_t ncand = *ncand_p;
    VT_SETUP
    if (c.perm && !fl.x) myers_loop<true>(ncand, c.perm, lens, words, b2, ref_length, e, c, vt);
    else myers_loop<false>(ncand, nullptr, lens, words, b2, ref_length, e, c, vt);
    vt.finish();
}

/* Two candidates per lane for e <= 7: the band (2e+1 <= 15 bits) fits a
 * 16-bit half, and a 16-bit band word gives the same distances and end
 * locations as the 32-bit one (the bits above the band converge the same
 * way; cf. the 32/64-bit equivalence), so candidate A runs in bits 0-15 and
 * B in bits 16-31 with packed 16-bit adds and shifts and shared bitwise ops.
 * Each half has its own read and window; both need the same read length. */
#define SWM_MAX_E_PACKED 7

typedef unsigned short swm_u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_add16(uint32_t a, uint32_t b)
{
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(swm_u16x2, a) + __builtin_bit_cast(swm_u16x2, b));
}
__device__ __forceinline__ uint32_t pk_shr16(uint32_t a)
{
    const swm_u16x2 one = {1, 1};
    return __builtin_bit_cast(uint32_t, __builtin_bit_cast(swm_u16x2, a) >> one);
}

__device__ __forceinline__ void myers_end_scan(uint32_t vp, uint32_t vn, int err, uint32_t L, uint32_t e,
                                               uint32_t ref_len, int32_t *err_out, int32_t *loc_out)
{
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

template <bool SHARED>
__device__ __forceinline__ void myers_pair(const uint32_t *__restrict__ rwa, const uint32_t *__restrict__ rwb,
                                           uint32_t L, uint32_t e, uint32_t rla, uint32_t rlb,
                                           uint32_t (&wla)[SWM_NW], uint32_t (&wha)[SWM_NW],
                                           uint32_t (&wlb)[SWM_NW], uint32_t (&whb)[SWM_NW], int32_t *ea,
                                           int32_t *la, int32_t *eb, int32_t *lb)
{
    uint32_t ra[16], rb[16];
#pragma unroll
    for (int j = 0; j < 16; j++) { ra[j] = rwa[j]; rb[j] = SHARED ? 0u : rwb[j]; }
    const uint32_t m16 = (1u << (2 * e + 1)) - 1u;
    const uint32_t mask = m16 | (m16 << 16);
    uint32_t vp = 0, vn = 0, ones = 0;
    const uint32_t nfull = L >> 5, rem = L & 31u;

#define SWM_PSTEP(S, RA, RB)                                                                \
    {                                                                                       \
        const int sb = 2 * ((S) & 15);                                                      \
        uint32_t RL, RH;                                                                    \
        if (SHARED) {                     /* one read: its base masks serve both halves */  \
            RL = (uint32_t)__builtin_amdgcn_sbfe((int)(RA), sb, 1);                         \
            RH = (uint32_t)__builtin_amdgcn_sbfe((int)(RA), sb + 1, 1);                     \
        } else {                                                                            \
            const uint32_t rla_ = (uint32_t)__builtin_amdgcn_sbfe((int)(RA), sb, 1);        \
            const uint32_t rha_ = (uint32_t)__builtin_amdgcn_sbfe((int)(RA), sb + 1, 1);    \
            const uint32_t rlb_ = (uint32_t)__builtin_amdgcn_sbfe((int)(RB), sb, 1);        \
            const uint32_t rhb_ = (uint32_t)__builtin_amdgcn_sbfe((int)(RB), sb + 1, 1);    \
            RL = (rla_ & 0xffffu) | (rlb_ & 0xffff0000u);                                   \
            RH = (rha_ & 0xffffu) | (rhb_ & 0xffff0000u);                                   \
        }                                                                                   \
        const uint32_t bl = __builtin_amdgcn_perm(__builtin_amdgcn_alignbit(lb1, lb0, (S)), \
                                                  __builtin_amdgcn_alignbit(la1, la0, (S)), 0x05040100u); \
        const uint32_t bh = __builtin_amdgcn_perm(__builtin_amdgcn_alignbit(hb1, hb0, (S)), \
                                                  __builtin_amdgcn_alignbit(ha1, ha0, (S)), 0x05040100u); \
        const uint32_t eq = ~((bl ^ RL) | (bh ^ RH)) & mask;                                \
        uint32_t x = eq | vn;                                                               \
        const uint32_t d0 = (pk_add16(x & vp, vp) ^ vp) | x;                                \
        const uint32_t hn = vp & d0;                                                        \
        const uint32_t hp = vn | ~(vp | d0);                                                \
        x = pk_shr16(d0);                                                                   \
        vn = x & hp;                                                                        \
        vp = hn | ~(x | hp);                                                                \
        ones = pk_add16(ones, d0 & 0x00010001u);                                            \
    }

    for (uint32_t cc = 0; cc < nfull; cc++) {
        const uint32_t a0 = ra[0], a1 = ra[1], b0 = rb[0], b1 = rb[1];
        const uint32_t la0 = wla[0], la1 = wla[1], ha0 = wha[0], ha1 = wha[1];
        const uint32_t lb0 = wlb[0], lb1 = wlb[1], hb0 = whb[0], hb1 = whb[1];
#pragma unroll
        for (int s = 0; s < 16; s++) SWM_PSTEP(s, a0, b0)
#pragma unroll
        for (int s = 16; s < 32; s++) SWM_PSTEP(s, a1, b1)
#pragma unroll
        for (int j = 0; j < SWM_NW - 1; j++) { wla[j] = wla[j + 1]; wha[j] = wha[j + 1]; wlb[j] = wlb[j + 1]; whb[j] = whb[j + 1]; }
#pragma unroll
        for (int j = 0; j < 14; j++) { ra[j] = ra[j + 2]; rb[j] = rb[j + 2]; }
    }
    if (rem) {
        const uint32_t a0 = ra[0], a1 = ra[1], b0 = rb[0], b1 = rb[1];
        const uint32_t la0 = wla[0], la1 = wla[1], ha0 = wha[0], ha1 = wha[1];
        const uint32_t lb0 = wlb[0], lb1 = wlb[1], hb0 = whb[0], hb1 = whb[1];
#pragma unroll
        for (int s = 0; s < 16; s++) {
            if ((uint32_t)s < rem) SWM_PSTEP(s, a0, b0)
        }
#pragma unroll
        for (int s = 16; s < 32; s++) {
            if ((uint32_t)s < rem) SWM_PSTEP(s, a1, b1)
        }
    }
#undef SWM_PSTEP
    myers_end_scan(vp & 0xffffu, vn & 0xffffu, (int)L - (int)(ones & 0xffffu), L, e, rla, ea, la);
    myers_end_scan(vp >> 16, vn >> 16, (int)L - (int)(ones >> 16), L, e, rlb, eb, lb);
}

__device__ __forceinline__ bool myers_one(uint32_t i, uint32_t st, uint32_t L, uint32_t ref_len,
                                          const uint32_t *__restrict__ words, const uint32_t *__restrict__ b2,
                                          uint32_t e, const CandSet &c)
{
    uint32_t wl[SWM_NW], wh[SWM_NW];
    load_window(b2, st, L, wl, wh);
    const uint32_t *w = words + (uint64_t)c.rs[i] * 32;
    int32_t er, lc;
    myers_lane<false>(w, L, e, ref_len, wl, wh, &er, &lc, nullptr, nullptr);
    c.err[i] = er;
    c.loc[i] = lc;
    return er <= (int32_t)e && ham_window(i, er, lc, st, L, w, b2, c);
}

/* The pair loop, once per order: with the identity a pair's loads do not
 * wait on an order lookup (one body for both orders had serialised them,
 * 2% of the skew step). */
template <bool PERM>
__device__ __forceinline__ void myers2_loop(uint32_t ncand, const uint32_t *__restrict__ perm,
                                            const uint32_t *__restrict__ lens, const uint32_t *__restrict__ words,
                                            const uint32_t *__restrict__ b2, uint64_t ref_length, uint32_t e,
                                            const CandSet &c, VerifyTally &vt)
{
    const uint32_t npair = (ncand + 1) / 2;
    for (uint32_t g0 = blockIdx.x * blockDim.x; g0 < npair; g0 += gridDim.x * blockDim.x) {
        const uint32_t g = g0 + threadIdx.x;
        uint32_t tq0 = ~0u, tq1 = ~0u;
        if (g < npair) {
            /* places 2g and 2g + 1 of the verification order */
            const bool has1 = 2 * g + 1 < ncand;
            const uint32_t i0 = PERM ? perm[2 * g] : 2 * g;
            const uint32_t i1 = PERM ? (has1 ? perm[2 * g + 1] : 0u) : 2 * g + 1;
            const uint32_t rs0 = c.rs[i0], rs1 = has1 ? c.rs[i1] : rs0;
            const uint32_t L0 = lens[rs0 >> 1], L1 = lens[rs1 >> 1];
            uint32_t rl0, rl1 = 0;
            const uint32_t st0 = pos_start(ref_length, e, c.val[i0], L0, &rl0);
            const uint32_t st1 = has1 ? pos_start(ref_length, e, c.val[i1], L1, &rl1) : 0xffffffffu;
            const bool v0 = st0 != 0xffffffffu, v1 = st1 != 0xffffffffu;
            if (!v0) { c.err[i0] = -2; c.loc[i0] = -1; vt.rej++; }
            if (has1 && !v1) { c.err[i1] = -2; c.loc[i1] = -1; vt.rej++; }
            vt.steps += (v0 ? L0 : 0u) + (v1 ? L1 : 0u);
            const bool packed = v0 && v1 && L0 == L1;
            /* the shared-read variant when every packed pair of the wave is
             * one (read, strand): wave-uniform, so no lane runs both */
            const bool shared = __all(!packed || rs0 == rs1);
            if (packed) {
                uint32_t wla[SWM_NW], wha[SWM_NW], wlb[SWM_NW], whb[SWM_NW];
                load_window(b2, st0, L0, wla, wha);
                load_window(b2, st1, L1, wlb, whb);
                const uint32_t *w0 = words + (uint64_t)rs0 * 32, *w1 = words + (uint64_t)rs1 * 32;
                int32_t e0, l0, e1, l1;
                if (shared)
                    myers_pair<true>(w0, w1, L0, e, rl0, rl1, wla, wha, wlb, whb, &e0, &l0, &e1, &l1);
                else
                    myers_pair<false>(w0, w1, L0, e, rl0, rl1, wla, wha, wlb, whb, &e0, &l0, &e1, &l1);
                c.err[i0] = e0; c.loc[i0] = l0;
                c.err[i1] = e1; c.loc[i1] = l1;
                if (e0 <= (int32_t)e && ham_window(i0, e0, l0, st0, L0, w0, b2, c)) tq0 = i0;
                if (e1 <= (int32_t)e && ham_window(i1, e1, l1, st1, L1, w1, b2, c)) tq1 = i1;
            } else {
                if (v0 && myers_one(i0, st0, L0, rl0, words, b2, e, c)) tq0 = i0;
                if (v1 && myers_one(i1, st1, L1, rl1, words, b2, e, c)) tq1 = i1;
            }
        }
        vt.round(tq0, tq1);
    }
}

__global__ void __launch_bounds__(256) k_myers2(const uint32_t *__restrict__ ncand_p, const uint32_t *__restrict__ lens,
                                                const uint32_t *__restrict__ words, const uint32_t *__restrict__ b2,
                                                uint64_t ref_length, uint32_t e, CandSet c, VT_ARGS)
{
    const uint2 fl = myers_flags(c);
    if (myers_stop(fl)) return;
    const uint32_t ncand = *ncand_p;
    VT_SETUP
    if (c.perm && !fl.x) myers2_loop<true>(ncand, c.perm, lens, words, b2, ref_length, e, c, vt);
    else myers2_loop<false>(ncand, nullptr, lens, words, b2, ref_length, e, c, vt);
    vt.finish();
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
    myers_lane<false>(read_words + (uint64_t)rr * 16, L, e, rlen[i], wl, wh, &er, &lc, nullptr, nullptr);
    err[i] = er;
    loc[i] = lc;
}

/* ----------------------------------------------------------------- k_trace */

/* The traceback of align_reads (:288-509) for one passing candidate whose
 * Hamming count differs from its edit distance: the record's start location
 * and the bytes its CIGAR adds (the latter drives the flush test).
 *
 * The reference keeps D0 and HP of every read base (d0_arry_64/hp_arry_64)
 * and walks them backwards.  Here the forward pass keeps (VP, VN) at every
 * 32-base chunk start; the walk recomputes one chunk at a time into
 * registers (static indices, no scratch) and walks it from its last base.
 * The CIGAR byte count is accumulated as the runs close: adjacent M/S runs
 * form one "<n>M" group, every I or D run its own "<n>I"/"<n>D". */
__device__ __forceinline__ void close_run(char c, int n, uint32_t &ms_acc, uint32_t &text, uint16_t *runs,
                                          uint32_t &nr)
{
    if (c == 'M' || c == 'S') {
        ms_acc += (uint32_t)n;
    } else {
        if (ms_acc) {
            text += (uint32_t)ndigits(ms_acc) + 1u;
            if (runs && nr < RUN_SLOT - 1) runs[1 + nr] = (uint16_t)(ms_acc | (RUN_M << 8));
            nr++;
            ms_acc = 0;
        }
        text += (uint32_t)ndigits((uint32_t)n) + 1u;
        if (runs && nr < RUN_SLOT - 1) runs[1 + nr] = (uint16_t)((uint32_t)n | ((c == 'I' ? RUN_I : RUN_D) << 8));
        nr++;
    }
}

/* RUNS: also store the merged runs (records path); a separate instance, so
 * the plain pipeline keeps the smaller register footprint. */
template <bool RUNS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_trace(const uint32_t *__restrict__ trace, const uint32_t *__restrict__ trace_count,
                                              const uint32_t *__restrict__ lens, const uint32_t *__restrict__ words,
                                              const uint32_t *__restrict__ b2, uint64_t ref_length, uint32_t e,
                                              CandSet c)
{
    __shared__ uint32_t s_ck[64 * 16];
    uint32_t *ck = s_ck + threadIdx.x * 16;
    const uint32_t nt = *trace_count;
    for (uint32_t ti = blockIdx.x * blockDim.x + threadIdx.x; ti < nt; ti += gridDim.x * blockDim.x) {
        const uint32_t ci = trace[ti];
        const uint32_t rs = c.rs[ci];
        const uint32_t L = lens[rs >> 1];
        uint32_t ref_len;
        const uint32_t st = pos_start(ref_length, e, c.val[ci], L, &ref_len);
        const uint32_t *w = words + (uint64_t)rs * 32;
        const uint32_t mask = (1u << (2 * e + 1)) - 1u;
        const uint32_t nfull = L >> 5, rem = L & 31u, nch = nfull + (rem ? 1u : 0u);

#define TR_STEP(S, R)                                                                      \
        {                                                                                  \
            const int sb = 2 * ((S) & 15);                                                 \
            uint32_t RL = (uint32_t)(((int32_t)((R) << (31 - sb))) >> 31);                 \
            uint32_t RH = (uint32_t)(((int32_t)((R) << (30 - sb))) >> 31);                 \
            uint32_t bl = __builtin_amdgcn_alignbit(l1, l0, (S));                          \
            uint32_t bh = __builtin_amdgcn_alignbit(h1, h0, (S));                          \
            uint32_t eq = ~((bl ^ RL) | (bh ^ RH)) & mask;                                 \
            uint32_t x = eq | vn;                                                          \
            uint32_t d0 = (((x & vp) + vp) ^ vp) | x;                                      \
            uint32_t hn = vp & d0;                                                         \
            uint32_t hp = vn | ~(vp | d0);                                                 \
            x = d0 >> 1;                                                                   \
            vn = x & hp;                                                                   \
            vp = hn | ~(x | hp);                                                           \
            d0r[(S)] = d0;                                                                 \
            hpr[(S)] = hp;                                                                 \
        }

        /* forward: checkpoints at chunk starts; each chunk's words are
         * loaded during the previous chunk's steps (its low window words are
         * the previous high ones; the last prefetch reads one word past the
         * read's window, inside the genome's zero padding) */
        uint32_t vp = 0, vn = 0;
        uint32_t d0r[32], hpr[32];      /* a chunk's D0 and HP (read by the backward pass) */
        uint32_t fr0 = w[0], fr1 = w[1];
        uint32_t fl0 = genome_word(b2, (int64_t)st, 0), fh0 = genome_word(b2, (int64_t)st, 1);
        uint32_t fl1 = genome_word(b2, (int64_t)st + 32, 0), fh1 = genome_word(b2, (int64_t)st + 32, 1);
        for (uint32_t cc = 0; cc < nfull; cc++) {
            ck[2 * cc] = vp;
            ck[2 * cc + 1] = vn;
            const uint32_t r0 = fr0, r1 = fr1, l0 = fl0, l1 = fl1, h0 = fh0, h1 = fh1;
            fr0 = w[2 * cc + 2];
            fr1 = w[2 * cc + 3];
            fl0 = l1;
            fh0 = h1;
            fl1 = genome_word(b2, (int64_t)st + 32 * cc + 64, 0);
            fh1 = genome_word(b2, (int64_t)st + 32 * cc + 64, 1);
#pragma unroll
            for (int s = 0; s < 16; s++) TR_STEP(s, r0)
#pragma unroll
            for (int s = 16; s < 32; s++) TR_STEP(s, r1)
        }
        if (rem) { ck[2 * nfull] = vp; ck[2 * nfull + 1] = vn; }

        /* backward walk (:363-466) */
        const int error = c.err[ci], location = c.loc[ci];
        int start_location = location - (int)L + 1;
        int match_site = location;
        int search_site = match_site - ((int)L - 1);
        int i = (int)L - 1;
        int sum_err = 0;
        bool first = true, done = false;
        char pre_char = 'N';
        int pre_size = 0;
        uint32_t ms_acc = 0, text = 2, nr = 0;
        uint16_t *runs = RUNS ? c.runs + (uint64_t)ci * RUN_SLOT : nullptr;
        for (int cc = (int)nch - 1; cc >= 0 && !done && i >= 0; cc--) {
            const uint32_t steps = ((uint32_t)cc == nfull) ? rem : 32u;
            const uint32_t r0 = w[2 * cc], r1 = w[2 * cc + 1];
            const int64_t g0 = (int64_t)st + 32 * (int64_t)cc;
            const uint32_t wml = genome_word(b2, g0 - 32, 0), wmh = genome_word(b2, g0 - 32, 1);
            const uint32_t l0 = genome_word(b2, g0, 0), l1 = genome_word(b2, g0 + 32, 0);
            const uint32_t h0 = genome_word(b2, g0, 1), h1 = genome_word(b2, g0 + 32, 1);
            /* The walk's three tests at a step s of this chunk are bits of
             * three masks over the chunk's 32 steps, all set by the search
             * site ss alone: D0 and HP at bit ss (0 outside the word), and
             * whether the window base at match_site (REF_LBIT/REF_HBIT,
             * :374-378) equals the read base -- match_site - i equals ss
             * (a match or substitution moves both, an insertion moves i
             * and ss, a deletion match_site and ss), so those window bases
             * are the 32 starting at chunk offset ss.  ss moves only on an
             * indel; then the masks are rebuilt from the chunk's D0/HP
             * registers.  A run of matches (D0 set, bases equal) is taken
             * in one step: each adds a base to the pending "M" run and
             * nothing else. */
            const uint32_t rdl = even_bits(r0) | (even_bits(r1) << 16);
            const uint32_t rdh = even_bits(r0 >> 1) | (even_bits(r1 >> 1) << 16);
            const bool far = e > 10u;                   /* ss can pass 31 */
            const uint32_t l2 = far ? genome_word(b2, g0 + 64, 0) : 0u, h2 = far ? genome_word(b2, g0 + 64, 1) : 0u;
            vp = ck[2 * cc];
            vn = ck[2 * cc + 1];
#pragma unroll
            for (int s = 0; s < 32; s++) {
                if ((uint32_t)s < steps) {
                    if (s < 16) TR_STEP(s, r0)
                    else TR_STEP(s, r1)
                }
            }
            /* Lanes walk until their search site moves or they leave the
             * chunk; then the wave rebuilds the masks of every lane that
             * moved at once (a rebuild is the wave's work, so one per
             * indel step of the slowest lane, not one per lane's). */
            uint32_t dm = 0, hm = 0, em = 0;
            bool stale = true, walking = !done && i >= 32 * cc;
            while (walking) {
                if (stale) {
                    stale = false;
                    dm = 0;
                    hm = 0;
                    if ((uint32_t)search_site < 32u) {
                        const uint32_t sh = (uint32_t)search_site;
#pragma unroll
                        for (int k = 0; k < 32; k++) {
                            dm |= ((d0r[k] >> sh) & 1u) << k;
                            hm |= ((hpr[k] >> sh) & 1u) << k;
                        }
                    }
                    uint32_t fl, fh;
                    if (search_site < -32 || search_site >= 64) {
                        fl = genome_word(b2, g0 + search_site, 0);
                        fh = genome_word(b2, g0 + search_site, 1);
                    } else {
                        const bool lo = search_site < 0, hi = search_site >= 32;
                        const uint32_t sh = (uint32_t)(search_site & 31);
                        fl = __builtin_amdgcn_alignbit(lo ? l0 : (hi ? l2 : l1), lo ? wml : (hi ? l1 : l0), sh);
                        fh = __builtin_amdgcn_alignbit(lo ? h0 : (hi ? h2 : h1), lo ? wmh : (hi ? h1 : h0), sh);
                    }
                    em = ~((fl ^ rdl) | (fh ^ rdh));
                }
                while (!stale) {
                    if (i < 32 * cc) { walking = false; break; }
                    if (!first && sum_err == error) { done = true; walking = false; break; }
                    const uint32_t sj = (uint32_t)(i - 32 * cc);
                    if (!first) {
                        /* matches from step sj down */
                        const uint32_t n = min((uint32_t)__clz(~(dm & em) << (31u - sj)), sj + 1u);
                        if (n) {
                            if (pre_char != 'M') { close_run(pre_char, pre_size, ms_acc, text, runs, nr); pre_char = 'M'; pre_size = 0; }
                            pre_size += (int)n;
                            i -= (int)n;
                            match_site -= (int)n;
                            continue;
                        }
                    }
                    const bool bd = (dm >> sj) & 1u, bh = (hm >> sj) & 1u, same = (em >> sj) & 1u;
                    /* a substitution closes like a match (both add to the
                     * pending "<n>M" group), so it takes the match's run
                     * label: only indels start a new run */
                    char nc;
                    if (bd && same) { i--; match_site--; nc = 'M'; }
                    else if (!bd && !same) { i--; match_site--; sum_err++; nc = 'M'; }
                    else if (bh) { i--; search_site++; sum_err++; start_location++; nc = first ? 'M' : 'I'; stale = true; }
                    else { search_site--; match_site--; sum_err++; start_location--; nc = 'D'; stale = true; }
                    if (first) { pre_char = nc; pre_size = 1; first = false; }
                    else if (nc != pre_char) { close_run(pre_char, pre_size, ms_acc, text, runs, nr); pre_char = nc; pre_size = 1; }
                    else pre_size++;
                }
            }
        }
#undef TR_STEP
        close_run(pre_char, pre_size, ms_acc, text, runs, nr);                       /* :468-473 */
        if (i >= 0) close_run('M', i + 1, ms_acc, text, runs, nr);
        if (ms_acc) {
            text += (uint32_t)ndigits(ms_acc) + 1u;
            if (runs && nr < RUN_SLOT - 1) runs[1 + nr] = (uint16_t)(ms_acc | (RUN_M << 8));
            nr++;
        }
        if (RUNS) runs[0] = (uint16_t)nr;
        const uint64_t new_loc = (uint64_t)st + (uint64_t)(int64_t)start_location + 1;
        c.rec[ci] = (uint32_t)new_loc;
        c.info[ci] = (text & 0xffffu) | INFO_TRACE | ((new_loc >> 32) ? INFO_NEG : 0u);
    }
}

/* ------------------------------------------------------------------ k_emit */

/* Emission of one read's records: its candidates in emission order (forward
 * strand, then reverse; batch by batch; groups of 8 valid lanes).  A Hamming
 * record is always written; a traceback record is skipped when it repeats
 * the previous record of the current block (:475-477).  After every group
 * the block is flushed when 8 more records or 256 more CIGAR bytes might not
 * fit (:1790, :1858), which resets the "previous record". */
struct Emit {
    uint32_t count, best_pos, best_dist, best_strand;
    uint32_t cigar_num, cigar_len, last_loc, cur_batch, vcount;
    bool open;
    __device__ void init()
    {
        count = 0; best_pos = 0xffffffffu; best_dist = 0xffu; best_strand = 0xffu;
    }
    __device__ void start_strand()
    {
        cigar_num = 0; cigar_len = 0; last_loc = 0; cur_batch = 0xffffffffu; vcount = 0; open = false;
    }
    __device__ void flush_check()
    {
        if (cigar_num + 8 > 256 || cigar_len + 16 * 16 > 256 * 16) { cigar_num = 0; cigar_len = 0; }
    }
    /* 0: the candidate's record is not written; WR_REC: written; WR_REC |
     * WR_BLOCK: written as the first record of an output block */
    __device__ uint64_t step(uint32_t bi, int32_t er, uint32_t rec, uint32_t info, uint32_t e, uint32_t strand)
    {
        const uint32_t b = bi >> 12;
        if (b != cur_batch) {
            if (open) { flush_check(); open = false; }
            cur_batch = b;
            vcount = 0;
        }
        if (er == -2) return 0;
        if (open && (vcount & 7u) == 0) { flush_check(); open = false; }
        open = true;
        vcount++;
        if (er > (int32_t)e) return 0;
        if ((info & INFO_TRACE) && cigar_num > 0 && !(info & INFO_NEG) && last_loc == rec) return 0;
        const uint64_t w = cigar_num == 0 ? (WR_REC | WR_BLOCK) : WR_REC;
        cigar_num++;
        cigar_len += info & 0xffffu;
        last_loc = rec;
        if (count == 0 || (uint32_t)er < best_dist) { best_pos = rec; best_dist = (uint32_t)er; best_strand = strand; }
        count++;
        return w;
    }
    __device__ void end_strand()
    {
        if (open) flush_check();
    }
    __device__ swm_read_result_t result() const
    {
        swm_read_result_t res;
        res.pos = count ? best_pos : 0xffffffffu;
        res.strand = count ? (uint8_t)best_strand : (uint8_t)0xff;
        res.dist = count ? (uint8_t)best_dist : (uint8_t)0xff;
        res.pad = 0;
        res.count = count;
        return res;
    }
};

__device__ __forceinline__ void record_partial(uint32_t count, uint32_t *part, uint32_t *s_sum)
{
    /* the workgroup's record count to its own slot: k_emit_combine adds the
     * slots up (one atomic per workgroup on one address cost 58 us of a
     * 1M-read step) */
    if (threadIdx.x == 0) *s_sum = 0;
    __syncthreads();
    if (count) atomicAdd(s_sum, count);
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = *s_sum;
}

#define EMIT_LOAD 4
__global__ void __launch_bounds__(128) k_emit(uint32_t nreads, uint32_t e, const uint32_t *__restrict__ cand_off,
                                              CandSet c, swm_read_result_t *__restrict__ out,
                                              uint32_t *__restrict__ rpart, uint32_t heavy_cands,
                                              uint32_t giant_cands, uint32_t *__restrict__ heavy,
                                              uint32_t *__restrict__ heavy_count)
{
    __shared__ uint32_t s_tmp[2], s_sum;
    if (*c.ovf) return;
    uint32_t rd = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = rd < nreads;
    if (!live) rd = 0;
    const uint32_t nc = cand_off[2 * rd + 2] - cand_off[2 * rd];
    const bool big = live && nc > heavy_cands;
    /* the largest reads first (heavy[0] .. [giant count)), the rest after
     * them from the list's other end: k_emit_heavy's waves start on the
     * slowest serial walks at once */
    const bool giant = big && nc > giant_cands;
    block_append(big && !giant, rd, heavy + nreads, heavy_count + 1, s_tmp);
    block_append(giant, rd, heavy, heavy_count, s_tmp);
    Emit em;
    em.init();
    if (live && !big) {
        for (uint32_t strand = 0; strand < 2; strand++) {
            const uint32_t rs = 2 * rd + strand;
            em.start_strand();
            const uint32_t cend = cand_off[rs + 1];
            /* four candidates' fields loaded together (a read has about
             * four), then stepped through in order */
            for (uint32_t c0 = cand_off[rs]; c0 < cend; c0 += EMIT_LOAD) {
                int32_t erq[EMIT_LOAD];
                uint32_t biq[EMIT_LOAD], recq[EMIT_LOAD], infoq[EMIT_LOAD];
#pragma unroll
                for (int k = 0; k < EMIT_LOAD; k++) {
                    const uint32_t ci = c0 + (uint32_t)k;
                    const bool in = ci < cend;
                    erq[k] = in ? c.err[ci] : -1;
                    biq[k] = in ? c.bi[ci] : 0u;
                    recq[k] = in ? c.rec[ci] : 0u;
                    infoq[k] = in ? c.info[ci] : 0u;
                }
#pragma unroll
                for (int k = 0; k < EMIT_LOAD; k++) {
                    const uint32_t ci = c0 + (uint32_t)k;
                    if (ci < cend) {
                        const int32_t er = erq[k];
                        const bool pass = er >= 0 && er <= (int32_t)e;
                        const uint32_t info = pass ? infoq[k] : 0u;
                        const uint64_t w = em.step(biq[k], er, pass ? recq[k] : 0u, info, e, strand);
                        if (c.wr) {
                            c.wr[ci] = w;
                            c.tb[ci] = w ? (info & 0xffffu) - 2u : 0u;
                        }
                    }
                }
            }
            em.end_strand();
        }
        out[rd] = em.result();
    }
    record_partial(live && !big ? em.count : 0u, rpart, &s_sum);
}

/* One wave per read with many candidates, 64 candidates per step in
 * parallel.  The emission state machine (Emit) reduces to two facts:
 *  - the "previous record" of the duplicate rule is always the record
 *    location of the previous passing candidate of the strand (a skipped
 *    duplicate has that location already), so a traceback record is a
 *    duplicate candidate ("dup") iff it is not negative and equals its
 *    passing predecessor; a dup is skipped unless no record has been written
 *    in the current output block;
 *  - the block counters only change at flush checks, which close each group
 *    of <= 8 valid lanes of a batch (:1790, :1858).
 * So the lanes classify themselves, and only the per-group counter update is
 * sequential: a dup that opens a group entered with an empty block is
 * written, every other dup is skipped, every other passing lane is written.
 * The best hit is the first minimum of (distance, strand, order) over the
 * written records. */
__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

/* One strand's share of a read's result: its record count and its best
 * record (key and location); the read's result combines the two. */
struct EmitPart {
    uint32_t count, best_rec;
    uint64_t best;
};

/* heavy[0 .. heavy_count[0]): giant reads; heavy[nreads ..
 * + heavy_count[1]): the other heavy reads.  A strand's emission state
 * starts afresh (the output block counters, the duplicate rule), so the
 * two strands of a read are separate work items -- slot 2h + strand, the
 * giant reads' first -- and a giant read's serial walk is split in two;
 * k_emit_combine then forms the read's result from the two parts. */
template <bool REC>   /* REC: the records path (c.wr set): also mark written records and block starts */
__global__ void __launch_bounds__(64) k_emit_heavy(const uint32_t *__restrict__ heavy,
                                                   const uint32_t *__restrict__ heavy_count, uint32_t nreads, uint32_t e,
                                                   const uint32_t *__restrict__ cand_off, CandSet c,
                                                   EmitPart *__restrict__ part, unsigned long long *n_records)
{
    const uint32_t lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t ng = heavy_count[0], nh = ng + heavy_count[1];
    for (uint32_t hs = blockIdx.x; hs < 2 * nh; hs += gridDim.x) {
        const uint32_t h = hs >> 1, strand = hs & 1u;
        const uint32_t rd = h < ng ? heavy[h] : heavy[nreads + (h - ng)];
        uint32_t count = 0, best_rec = 0;
        uint64_t best = ~0ull;
        {
            const uint32_t rs = 2 * rd + strand;
            const uint32_t c0 = cand_off[rs], cend = cand_off[rs + 1];
            /* the longest walks set the kernel's length and share their SIMD
             * with up to 7 other waves: they issue first */
            if (cend - c0 > 8u * GIANT_CANDS) __builtin_amdgcn_s_setprio(3);
            else if (cend - c0 > 2u * GIANT_CANDS) __builtin_amdgcn_s_setprio(2);
            else if (cend - c0 > GIANT_CANDS / 2u) __builtin_amdgcn_s_setprio(1);
            else __builtin_amdgcn_s_setprio(0);
            /* carries between windows */
            uint32_t cur_b = 0xffffffffu, vcnt = 0, prevrec = 0;
            bool has_prev = false;
            uint32_t num = 0, len = 0;          /* output block counters */
            /* the open group */
            bool g_open = false, g_seen = false, g_firstd = false, g_wseen = false;
            uint32_t g_n = 0, g_len = 0, g_ftext = 0, g_frec = 0, g_fci = 0, g_wci = 0;
            uint64_t g_fkey = 0;
            auto close_group = [&]() {
                if (!g_open) return;
                if (num == 0 && g_firstd) {
                    num = 1;
                    len = g_ftext;
                    count++;
                    if (g_fkey < best) { best = g_fkey; best_rec = g_frec; }
                    /* the lane that stored this candidate's "not written" flag
                     * overwrites it; the record opens the output block */
                    if (REC && lane == ((g_fci - c0) & 63u)) {
                        c.wr[g_fci] = WR_REC | WR_BLOCK;
                        c.tb[g_fci] = g_ftext - 2u;
                    }
                } else if (REC && num == 0 && g_wseen && lane == ((g_wci - c0) & 63u)) {
                    /* the group's first plainly written record opens the block */
                    c.wr[g_wci] = WR_REC | WR_BLOCK;
                }
                num += g_n;
                len += g_len;
                if (num + 8 > 256 || len + 16 * 16 > 256 * 16) { num = 0; len = 0; }
                g_open = false;
            };
            /* the four fields of a window are loaded together, four windows
             * ahead: a giant read is one wave's serial walk, so the loads
             * must not wait on the walk.  Four named slots, rotated by
             * position (no loaded register is ever moved). */
            auto fetch = [&](uint32_t at, int32_t &fe, uint32_t &fb, uint32_t &fr, uint32_t &fi) {
                const uint32_t ci = at + lane;
                const bool ok = ci < cend;
                const uint32_t cc = ok ? ci : c0;             /* in-range address; value discarded */
                const int32_t le = c.err[cc];
                const uint32_t lb = c.bi[cc], lr = c.rec[cc], li = c.info[cc];
                fe = ok ? le : -2;
                fb = ok ? lb : 0xffffffffu << 12;
                fr = ok ? lr : 0u;
                fi = ok ? li : 0u;
            };
/* one window of 64 candidates, its fields already loaded */
            auto window = [&](uint32_t base, int32_t f_er, uint32_t f_bi, uint32_t f_rec, uint32_t f_info) {
                const uint32_t ci = base + lane;
                const uint32_t m = min(64u, cend - base);
                const bool in = ci < cend;
                const int32_t er = f_er;
                const uint32_t b = in ? (f_bi >> 12) : 0xffffffffu;
                const bool valid = er != -2;
                const bool pass = valid && er <= (int32_t)e;
                const uint32_t rec = pass ? f_rec : 0u;
                const uint32_t info = pass ? f_info : 0u;
                /* duplicate candidates: predecessor = last passing lane below, else the carry */
                const uint64_t pm = __ballot(pass);
                const uint64_t pb = pm & below;
                const uint32_t src = pb ? 63u - (uint32_t)__clzll(pb) : lane;
                const uint32_t prv_l = (uint32_t)__shfl((int)rec, (int)src, 64);
                const bool hp = pb ? true : has_prev;
                const uint32_t prv = pb ? prv_l : prevrec;
                const bool dup = pass && (info & INFO_TRACE) && !(info & INFO_NEG) && hp && rec == prv;
                /* index of the lane among the valid lanes of its batch; the
                 * previous lane's batch by wave_shr:1 (lane 0 takes the carry) */
                const uint32_t bprev = (uint32_t)__builtin_amdgcn_update_dpp((int)cur_b, (int)b, 0x138, 0xf, 0xf, false);
                const uint64_t sm = __ballot(in && b != bprev);
                const uint64_t vm = __ballot(valid);
                const uint64_t smb = sm & (below | (1ull << lane));
                uint32_t v;
                if (smb) {
                    const uint32_t s0 = 63u - (uint32_t)__clzll(smb);
                    v = (uint32_t)__popcll(vm & below & ~((1ull << s0) - 1ull));
                } else {
                    v = vcnt + (uint32_t)__popcll(vm & below);
                }
                const uint64_t gm = __ballot(valid && (v & 7u) == 0);
                /* plainly written lanes */
                const bool wr = pass && !dup;
                if (REC && in) {
                    c.wr[ci] = wr ? WR_REC : 0ull;
                    c.tb[ci] = wr ? (info & 0xffffu) - 2u : 0u;
                }
                const uint64_t am = __ballot(wr);
                const uint32_t tinc = wave_incl_scan_dpp(wr ? (info & 0xffffu) : 0u);
                /* best hit: every earlier record (written or pending) precedes
                 * this window, so a lane here beats it only with a smaller
                 * distance; take the first lane of the smallest one */
                uint32_t best_er = best == ~0ull ? 0xffffffffu : (uint32_t)(best >> 40);
                for (uint64_t cm = __ballot(wr && (uint32_t)er < best_er); cm;
                     cm = __ballot(wr && (uint32_t)er < best_er)) {
                    const uint32_t f = (uint32_t)__ffsll((unsigned long long)cm) - 1u;
                    best_er = rdlane((uint32_t)er, f);
                    best = ((uint64_t)best_er << 40) | ((uint64_t)strand << 32) | (uint64_t)(base + f - c0);
                    best_rec = rdlane(rec, f);
                }
                count += (uint32_t)__popcll(am);
                /* Segments: lane 0 (the open group's continuation) and every
                 * group start.  The lane of a segment start computes the
                 * segment's written count, CIGAR bytes and first passing
                 * lane, so the serial part reads a few scalars per group. */
                const uint64_t segs = gm | 1ull;
                const uint64_t above = segs & ~((2ull << lane) - 1ull);
                const uint32_t send = above ? min((uint32_t)__ffsll((unsigned long long)above) - 1u, m) : m;
                const uint64_t srm = (send >= 64u ? ~0ull : ((1ull << send) - 1ull)) & ~((1ull << lane) - 1ull);
                const uint32_t t_end = (uint32_t)__shfl((int)tinc, (int)(send ? send - 1u : 0u), 64);
                const uint32_t t_beg = tinc - (wr ? (info & 0xffffu) : 0u);
                const uint32_t s_len = send > lane ? t_end - t_beg : 0u;
                const uint64_t spr = pm & srm;
                const uint32_t sf = spr ? (uint32_t)__ffsll((unsigned long long)spr) - 1u : 0u;
                const uint64_t dm = __ballot(dup);
                const uint64_t wsm = am & srm;    /* the segment's plainly written lanes */
                uint32_t s_pack = (uint32_t)__popcll(wsm) | (spr ? 0x100u : 0u) |
                                  ((spr && ((dm >> sf) & 1ull)) ? 0x200u : 0u);
                if (REC) {
                    const uint32_t sw = wsm ? (uint32_t)__ffsll((unsigned long long)wsm) - 1u : 0u;
                    s_pack |= (wsm ? 0x400u : 0u) | (sw << 16);
                }
                const uint32_t q_info = (uint32_t)__shfl((int)info, (int)sf, 64);
                const uint32_t q_rec = (uint32_t)__shfl((int)rec, (int)sf, 64);
                const uint32_t q_er = (uint32_t)__shfl((int)er, (int)sf, 64);
                for (uint64_t rest = segs; rest; rest &= rest - 1ull) {
                    const uint32_t p = (uint32_t)__ffsll((unsigned long long)rest) - 1u;
                    if ((gm >> p) & 1ull) {
                        close_group();
                        g_open = true; g_seen = false; g_firstd = false; g_wseen = false; g_n = 0; g_len = 0;
                    }
                    if (!g_open) continue;
                    const uint32_t pk = rdlane(s_pack, p);
                    if (REC && !g_wseen && (pk & 0x400u)) { g_wseen = true; g_wci = base + (pk >> 16); }
                    g_n += pk & 0xffu;
                    g_len += rdlane(s_len, p);
                    if (!g_seen && (pk & 0x100u)) {
                        g_seen = true;
                        if (pk & 0x200u) {
                            const uint32_t f = rdlane(sf, p);
                            g_firstd = true;
                            g_ftext = rdlane(q_info, p) & 0xffffu;
                            g_frec = rdlane(q_rec, p);
                            g_fci = base + f;
                            g_fkey = ((uint64_t)rdlane(q_er, p) << 40) | ((uint64_t)strand << 32) |
                                     (uint64_t)(base + f - c0);
                        }
                    }
                }
                /* carries */
                const uint32_t last = m - 1u;
                cur_b = rdlane(b, last);
                vcnt = rdlane(v, last) + (uint32_t)((vm >> last) & 1ull);
                if (pm) { has_prev = true; prevrec = rdlane(rec, 63u - (uint32_t)__clzll(pm)); }
            };
            int32_t e0, e1, e2, e3, e4, e5, e6, e7;
            uint32_t b0, b1, b2, b3, b4, b5, b6, b7, r0, r1, r2, r3, r4, r5, r6, r7;
            uint32_t i0, i1, i2, i3, i4, i5, i6, i7;
            fetch(c0, e0, b0, r0, i0);
            fetch(c0 + 64u, e1, b1, r1, i1);
            fetch(c0 + 128u, e2, b2, r2, i2);
            fetch(c0 + 192u, e3, b3, r3, i3);
            fetch(c0 + 256u, e4, b4, r4, i4);
            fetch(c0 + 320u, e5, b5, r5, i5);
            fetch(c0 + 384u, e6, b6, r6, i6);
            fetch(c0 + 448u, e7, b7, r7, i7);
#define EMIT_SLOT(E, B, R, I, OFF)                                                  \
            if (base + (OFF) < cend) {                                              \
                const int32_t fe = E;                                               \
                const uint32_t fb = B, fr = R, fi = I;                              \
                fetch(base + (OFF) + 512u, E, B, R, I);                             \
                window(base + (OFF), fe, fb, fr, fi);                               \
            }
            for (uint32_t base = c0; base < cend; base += 512u) {
                EMIT_SLOT(e0, b0, r0, i0, 0u)
                EMIT_SLOT(e1, b1, r1, i1, 64u)
                EMIT_SLOT(e2, b2, r2, i2, 128u)
                EMIT_SLOT(e3, b3, r3, i3, 192u)
                EMIT_SLOT(e4, b4, r4, i4, 256u)
                EMIT_SLOT(e5, b5, r5, i5, 320u)
                EMIT_SLOT(e6, b6, r6, i6, 384u)
                EMIT_SLOT(e7, b7, r7, i7, 448u)
            }
#undef EMIT_SLOT
            close_group();
        }
        if (lane == 0) {
            EmitPart pt;
            pt.count = count;
            pt.best_rec = best_rec;
            pt.best = best;
            part[hs] = pt;
            if (count) atomicAdd(n_records, (unsigned long long)count);
        }
    }
}

/* A heavy read's result from its strands' parts: the counts add up and the
 * best record is the smaller key (distance, strand, order).  Workgroups
 * 0..7 also add up k_emit's per-workgroup record counts. */
__global__ void __launch_bounds__(256) k_emit_combine(const uint32_t *__restrict__ heavy,
                                                     const uint32_t *__restrict__ heavy_count, uint32_t nreads,
                                                     const EmitPart *__restrict__ part,
                                                     swm_read_result_t *__restrict__ out,
                                                     const uint32_t *__restrict__ rpart, uint32_t nrpart,
                                                     unsigned long long *n_records)
{
    if (blockIdx.x < 8) {
        __shared__ unsigned long long s_sum[1];
        if (threadIdx.x == 0) s_sum[0] = 0;
        __syncthreads();
        unsigned long long t = 0;
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nrpart; k += 8 * blockDim.x) t += rpart[k];
        if (t) atomicAdd(&s_sum[0], t);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (s_sum[0]) atomicAdd(n_records, s_sum[0]);
        }
    }
    const uint32_t ng = heavy_count[0], nh = ng + heavy_count[1];
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += gridDim.x * blockDim.x) {
        const uint32_t rd = h < ng ? heavy[h] : heavy[nreads + (h - ng)];
        const EmitPart a = part[2 * h], b = part[2 * h + 1];
        const uint32_t count = a.count + b.count;
        const bool second = b.best < a.best;
        const uint64_t best = second ? b.best : a.best;
        swm_read_result_t res;
        res.pos = count ? (second ? b.best_rec : a.best_rec) : 0xffffffffu;
        res.strand = count ? (uint8_t)((best >> 32) & 1ull) : (uint8_t)0xff;
        res.dist = count ? (uint8_t)(best >> 40) : (uint8_t)0xff;
        res.pad = 0;
        res.count = count;
        out[rd] = res;
    }
}

/* ---------------------------------------------------------------- k_records */

/* Decimal digits of v (none for 0, as sprint_length, mapper_cpe.c:214-228). */
__device__ __forceinline__ uint32_t put_dec(char *o, uint32_t v)
{
    const uint32_t n = (uint32_t)ndigits(v);
    for (uint32_t k = n; k > 0; k--) {
        o[k - 1] = (char)('0' + v % 10u);
        v /= 10u;
    }
    return n;
}

/* Every written record at its rank among the written candidates (ridx, the
 * exclusive scan of wr: low half) with its CIGAR text at toff (the scan of
 * tb): "<L>M" for a Hamming record (mapper_cpe.c:283-287), else the
 * traceback's merged runs from the last recorded to the first (:479-507).  A
 * record is primary when no candidate of its (read, strand) before it was
 * written (main_flag, :1462-1465).  A record that opens an output block (the
 * first written after a put_cigar flush, :1461-1500) gives its rank to the
 * block's entry (ridx's high half is the block's rank). */
__global__ void __launch_bounds__(256) k_records(const uint32_t *__restrict__ ncand_p, const uint32_t *__restrict__ lens,
                                                 const uint32_t *__restrict__ cand_off, CandSet c,
                                                 const uint64_t *__restrict__ ridx, const uint32_t *__restrict__ toff,
                                                 swm_record_t *__restrict__ out, char *__restrict__ text,
                                                 uint32_t *__restrict__ block_first)
{
    const uint32_t ncand = *ncand_p;
    for (uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x; ci < ncand; ci += gridDim.x * blockDim.x) {
        const uint64_t w = c.wr[ci];
        if (!w) continue;
        const uint32_t rs = c.rs[ci];
        const uint64_t rb = ridx[ci];
        const uint32_t r = (uint32_t)rb;
        if (w & WR_BLOCK) block_first[rb >> 32] = r;
        const uint32_t off = toff[ci];
        char *o = text + off;
        uint32_t len = 0;
        if (c.info[ci] & INFO_TRACE) {
            const uint16_t *ru = c.runs + (uint64_t)ci * RUN_SLOT;
            const uint32_t nr = min((uint32_t)ru[0], (uint32_t)RUN_SLOT - 1u);
            for (uint32_t k = nr; k > 0; k--) {
                const uint32_t v = ru[k];
                len += put_dec(o + len, v & 0xffu);
                const uint32_t kind = v >> 8;
                o[len++] = kind == RUN_M ? 'M' : (kind == RUN_I ? 'I' : 'D');
            }
        } else {
            len = put_dec(o, lens[rs >> 1]);
            o[len++] = 'M';
        }
        swm_record_t x;
        x.read = rs >> 1;
        x.pos = c.rec[ci];
        x.cigar_off = off;
        x.strand = (uint8_t)(rs & 1u);
        x.dist = (uint8_t)c.err[ci];
        x.primary = (uint32_t)ridx[cand_off[rs]] == r ? 1 : 0;
        x.cigar_len = (uint8_t)len;
        out[r] = x;
    }
}

/* Reads by candidate count (both strands) in log2 bins, and the maximum. */
__global__ void __launch_bounds__(256) k_hist(uint32_t nreads, const uint32_t *__restrict__ cand_off,
                                              unsigned long long *__restrict__ hist)
{
    __shared__ uint32_t bins[25];
    if (threadIdx.x < 25) bins[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t rd = blockIdx.x * blockDim.x + threadIdx.x; rd < nreads; rd += gridDim.x * blockDim.x) {
        const uint32_t c = cand_off[2 * rd + 2] - cand_off[2 * rd];
        const uint32_t b = c ? min(32u - (uint32_t)__clz(c), 23u) : 0u;
        atomicAdd(&bins[b], 1u);
        atomicMax(&bins[24], c);
    }
    __syncthreads();
    if (threadIdx.x < 24 && bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)bins[threadIdx.x]);
    if (threadIdx.x == 24) atomicMax(&hist[24], (unsigned long long)bins[24]);
}

/* ------------------------------------------------------- scans and checks */

/* Exclusive prefix sums in two launches (tile sums; tile rescan, each tile
 * adding up the sums of the tiles before it, tile 0 also the total and its
 * capacity check), graph-safe: plain kernels, no memset and no
 * temporary-size query. */
#define SCAN_NT 256
#define SCAN_PER 16
#define SCAN_TILE (SCAN_NT * SCAN_PER)

/* One scan job covers up to two arrays of the same length (blockIdx.y picks
 * the array): array 0 is u64, array 1 is u32; tile sums are u64 for both.
 * The rescan's first tile also checks each total against its capacity
 * (setting the overflow flag) and keeps it in the counters. */
struct ScanJob {
    const uint64_t *in0;
    uint64_t *out0;
    const uint32_t *in1;
    uint32_t *out1;
    uint64_t n;
    uint64_t *sums;          /* [2][nb] */
    uint32_t nb;
    uint32_t first;          /* first array (0 or 1); blockIdx.y counts from it */
    uint64_t cap[2];
    unsigned long long *tot[2];
    uint32_t *ovf;
    /* with cls_nsel, the scan of the position counts (array 0) also sorts
     * the heavy items (cls_n of them) into the seed-heap kernels' lists */
    const uint32_t *cls_nsel;
    uint32_t cls_n;
    uint32_t *cls_heavy, *cls_heavy_count, *cls_giant, *cls_giant_count;
    uint32_t giant_pos;      /* items with more positions go to the giant list */
    uint32_t cls_heavy_on;   /* heavy and giant lists (the work-queue form) */
    uint32_t *cls_irreg, *cls_irreg_count;   /* items with an irregular bucket (every form) */
};

/* The SCAN_PER values of thread t in a tile (elements q * SCAN_NT + t),
 * loaded unconditionally (a place past the end reads element 0 and counts
 * 0), so that all of them are in flight at once: a bounds test per load had
 * made each load wait for the one before. */
template <typename T>
__device__ __forceinline__ void scan_tile_load(const T *__restrict__ in, uint64_t n, uint64_t tile,
                                               uint64_t (&x)[SCAN_PER])
{
    T raw[SCAN_PER];
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const uint64_t g = tile + (uint64_t)q * SCAN_NT + threadIdx.x;
        raw[q] = in[g < n ? g : 0u];
    }
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const uint64_t g = tile + (uint64_t)q * SCAN_NT + threadIdx.x;
        x[q] = g < n ? (uint64_t)raw[q] : 0ull;
    }
}

__global__ void __launch_bounds__(SCAN_NT) k_scan_reduce(ScanJob j)
{
    __shared__ uint64_t s_w[SCAN_NT / 64];
    const uint32_t y = j.first + blockIdx.y;
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t x[SCAN_PER];
    if (y == 0) scan_tile_load(j.in0, j.n, base, x);
    else scan_tile_load(j.in1, j.n, base, x);
    uint64_t v = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) v += x[q];
    uint64_t tot;
    (void)block_incl_scan<uint64_t, SCAN_NT>(v, s_w, &tot);
    if (threadIdx.x == 0) j.sums[(uint64_t)y * j.nb + blockIdx.x] = tot;
}

/* LDS index of tile element e: one pad word per 16, so that a thread's 16
 * consecutive elements (17 t + q) do not all sit in one bank */
#define SCAN_PAD(e) ((e) + ((e) >> 4))

__global__ void __launch_bounds__(SCAN_NT) k_scan_down(ScanJob j)
{
    __shared__ uint64_t s_w[SCAN_NT / 64];
    __shared__ uint32_t s_base[3];
    __shared__ uint64_t s_v[SCAN_TILE + SCAN_TILE / 16];
    const uint32_t y = j.first + blockIdx.y;
    const uint32_t tid = threadIdx.x;
    const uint64_t tile = (uint64_t)blockIdx.x * SCAN_TILE;
    const uint64_t *sums = j.sums + (uint64_t)y * j.nb;
    /* this tile's offset: the tile sums before it (tile 0: all of them,
     * for the total) */
    uint64_t pre = 0, all = 0;
    for (uint32_t b = tid; b < j.nb; b += SCAN_NT) {
        const uint64_t x = sums[b];
        if (b < blockIdx.x) pre += x;
        all += x;
    }
    uint64_t tile_off, grand;
    (void)block_incl_scan<uint64_t, SCAN_NT>(pre, s_w, &tile_off);
    (void)block_incl_scan<uint64_t, SCAN_NT>(blockIdx.x == 0 ? all : 0ull, s_w, &grand);
    if (blockIdx.x == 0 && tid == 0) {
        if (j.tot[y]) *j.tot[y] = grand;
        if (grand > j.cap[y]) *j.ovf = 1u;
    }
    /* the tile in and out through LDS, thread t on elements q * SCAN_NT + t
     * (coalesced; with each thread on 16 consecutive elements every wave
     * access touched 64 lines); the heavy items are classified on the way
     * in (heavy: a selected 12-mer with more than one chunk; over giant_pos
     * positions to the giant list) */
    const bool cls = y == 0 && j.cls_nsel;
    uint64_t xv[SCAN_PER];
    if (y == 0) scan_tile_load(j.in0, j.n, tile, xv);
    else scan_tile_load(j.in1, j.n, tile, xv);
    uint32_t hm = 0, gm = 0, im = 0;
    if (cls) {
        uint32_t ns[SCAN_PER];
#pragma unroll
        for (int q = 0; q < SCAN_PER; q++) {
            const uint64_t g = tile + (uint64_t)q * SCAN_NT + tid;
            ns[q] = j.cls_nsel[g < j.cls_n ? g : 0u];
        }
#pragma unroll
        for (int q = 0; q < SCAN_PER; q++) {
            const uint64_t g = tile + (uint64_t)q * SCAN_NT + tid;
            if (g >= j.cls_n) continue;
            if (j.cls_heavy_on && (ns[q] & NSEL_HEAVY)) {
                /* heavy items with an irregular bucket too: their batches
                 * with a descending chunk go to k_merge */
                if (xv[q] > j.giant_pos) gm |= 1u << q;
                else hm |= 1u << q;
            } else if (ns[q] & NSEL_IRREG) {
                im |= 1u << q;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) s_v[SCAN_PAD((uint32_t)q * SCAN_NT + tid)] = xv[q];
    __syncthreads();
    uint64_t v[SCAN_PER];
    uint64_t t = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        v[q] = s_v[SCAN_PAD(tid * SCAN_PER + (uint32_t)q)];
        t += v[q];
    }
    uint64_t tot;
    uint64_t run = block_incl_scan<uint64_t, SCAN_NT>(t, s_w, &tot) - t + tile_off;   /* its barriers end the reads */
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        s_v[SCAN_PAD(tid * SCAN_PER + (uint32_t)q)] = run;
        run += v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < SCAN_PER; q++) {
        const uint32_t e = (uint32_t)q * SCAN_NT + tid;
        const uint64_t g = tile + e;
        if (g < j.n) {
            if (y == 0) j.out0[g] = s_v[SCAN_PAD(e)];
            else j.out1[g] = (uint32_t)s_v[SCAN_PAD(e)];
        }
    }
    if (cls) {
        /* one atomic per list per workgroup; the three counts of a thread
         * (at most SCAN_TILE each) in 21-bit fields of one scanned word */
        const uint64_t c = (uint64_t)__popc(hm) | ((uint64_t)__popc(gm) << 21) | ((uint64_t)__popc(im) << 42);
        uint64_t ctot;
        const uint64_t cex = block_incl_scan<uint64_t, SCAN_NT>(c, s_w, &ctot) - c;
        const uint32_t th = (uint32_t)(ctot & 0x1fffffu), tg = (uint32_t)((ctot >> 21) & 0x1fffffu),
                       ti = (uint32_t)(ctot >> 42);
        if (tid == 0) {
            s_base[0] = th ? atomicAdd(j.cls_heavy_count, th) : 0u;
            s_base[1] = tg ? atomicAdd(j.cls_giant_count, tg) : 0u;
            s_base[2] = ti ? atomicAdd(j.cls_irreg_count, ti) : 0u;
        }
        __syncthreads();
        uint32_t ph = s_base[0] + (uint32_t)(cex & 0x1fffffu), pg = s_base[1] + (uint32_t)((cex >> 21) & 0x1fffffu),
                 pi = s_base[2] + (uint32_t)(cex >> 42);
#pragma unroll
        for (int q = 0; q < SCAN_PER; q++) {
            const uint32_t g = (uint32_t)(tile + (uint64_t)q * SCAN_NT + tid);
            if ((hm >> q) & 1u) j.cls_heavy[ph++] = g;
            if ((gm >> q) & 1u) j.cls_giant[pg++] = g;
            if ((im >> q) & 1u) j.cls_irreg[pi++] = g;
        }
    }
}

/* --------------------------------------------------------------- host side */

inline uint32_t blocks_for(uint64_t n, uint32_t bs) { return (uint32_t)((n + bs - 1) / bs); }

}  // namespace

/* Device workspace of a slot, grown on demand and reused by every batch (no
 * allocation on the steady path).  gen changes whenever a buffer moves, which
 * retires the graphs captured over the old addresses. */
struct swm_ws {
    struct Buf { void *p = nullptr; size_t cap = 0; };
    enum { WORDS, SEL, SELC, NSEL, NPOS, NCHUNK, SURV_OFF, BD_OFF, SURV, BDESC, NBATCH, NUNIQ, CAND_OFF, C_VAL, C_RS, C_BI,
           C_ERR, C_LOC, C_REC, C_INFO, COUNTERS, BIG, TRACE, HEAVY, HEAVY_C, HEAVY_E, SCAN_TMP, IN_READS, IN_LENS,
           OUT, HIST, R_WR, R_TB, R_RUNS, R_IDX, R_TOFF, R_SCAN, R_TOT, R_OUT, R_TEXT, R_BLOCK, CDESC, HBATCH, GIANT, EPART, RPART, LBINS, PERM, IRREG, MERGE, NBUF };
    Buf b[NBUF];
    uint64_t gen = 1;
    ~swm_ws() { for (auto &x : b) if (x.p) (void)hipFree(x.p); }
    template <typename T>
    int get(int id, T **out, size_t n)
    {
        size_t bytes = (n ? n : 1) * sizeof(T);
        if (b[id].cap < bytes) {
            if (b[id].p) (void)hipFree(b[id].p);
            b[id].p = nullptr;
            b[id].cap = 0;
            gen++;
            size_t want = bytes + bytes / 4;
            if (hipMalloc(&b[id].p, want) != hipSuccess) return SWM_ENOMEM;
            b[id].cap = want;
        }
        *out = (T *)b[id].p;
        return SWM_OK;
    }
    uint64_t bytes() const
    {
        uint64_t t = 0;
        for (const auto &x : b) t += x.cap;
        return t;
    }
};

/* captured graphs kept per slot: at least the serving loop's launch sizes
 * (powers of two from 1024 to 2^18 reads, nine), so that after its warmup
 * no launch captures */
#define SLOT_GRAPHS 16

/* One stream with its own workspace: batches on different slots run
 * concurrently over the shared, read-only index. */
struct swm_slot {
    hipStream_t stream = nullptr;
    /* second stream of the slot: the heavy and irregular plan kernels run on
     * it beside the light-item kernel (forked and joined with events, also
     * inside a captured graph) */
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    swm_ws ws;
    swm_stats_t stats{};
    /* largest observed per-read totals: they size the fixed capacities */
    double r_pos = 0, r_chunk = 0, r_cand = 0;
    struct GraphEnt {
        bool used = false;
        const void *reads = nullptr, *lens = nullptr, *out = nullptr;
        uint64_t n = 0, stride = 0, gen = 0, tick = 0;
        swm_opts opt{};
        uint32_t e = 0;
        hipGraphExec_t exec = nullptr;
    };
    GraphEnt g[SLOT_GRAPHS];
    uint64_t tick = 0;
    hipEvent_t ev[8] = {};      /* stage bounds [0..5]; k_plan_pair [6], [7] */
    /* page-locked staging, so that no copy goes through the runtime's shared
     * pageable-memory path (which serialises the slots): totals read back
     * between stages, counters, and host batches in and out */
    uint64_t *h_small = nullptr;     /* [0..11] counters at the end; before that [8] positions, [9] chunks,
                                        [10] candidates, [11..13] record totals */
    struct Pin { void *p = nullptr; size_t cap = 0; };
    Pin h_reads, h_lens, h_out;
    double t_entry = 0;   /* host clock at the start of the current call */
    /* reads the workspace is sized for from the first graph capture on, so
     * that smaller batches never move a buffer (which would retire every
     * captured graph of the slot) */
    uint64_t reserve_n = 0;
    swm_gate *gate = nullptr;         /* shared with the other slots of a server */
    hipEvent_t ev_gate = nullptr;     /* the end of this slot's last run */
    ~swm_slot()
    {
        if (ev_gate) (void)hipEventDestroy(ev_gate);
        if (stream) (void)hipStreamSynchronize(stream);
        if (side) (void)hipStreamSynchronize(side);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        for (auto &x : g) if (x.exec) (void)hipGraphExecDestroy(x.exec);
        for (auto &x : ev) if (x) (void)hipEventDestroy(x);
        for (Pin *x : {&h_reads, &h_lens, &h_out}) if (x->p) (void)hipHostFree(x->p);
        if (h_small) (void)hipHostFree(h_small);
        if (side) (void)hipStreamDestroy(side);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int pin(Pin &x, size_t bytes)
    {
        if (x.cap >= bytes) return SWM_OK;
        if (x.p) (void)hipHostFree(x.p);
        x.p = nullptr;
        x.cap = 0;
        const size_t want = bytes + bytes / 4;
        if (hipHostMalloc(&x.p, want, hipHostMallocDefault) != hipSuccess) return SWM_ENOMEM;
        x.cap = want;
        return SWM_OK;
    }
};

swm_slot *swm_slot_create(swm_ctx_t *ctx)
{
    (void)ctx;
    swm_slot *s = new (std::nothrow) swm_slot();
    if (!s) return nullptr;
    bool ok = hipStreamCreate(&s->stream) == hipSuccess && hipStreamCreate(&s->side) == hipSuccess;
    for (auto &x : s->ev) ok = ok && hipEventCreate(&x) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&s->ev_gate, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&s->h_small, 16 * sizeof(uint64_t), hipHostMallocDefault) == hipSuccess;
    if (!ok) { delete s; return nullptr; }
    return s;
}

void swm_slot_destroy(swm_slot *s) { delete s; }

const swm_stats_t *swm_slot_stats(const swm_slot *s) { return &s->stats; }

uint64_t swm_slot_bytes(const swm_slot *s) { return s->ws.bytes(); }

void swm_slot_reserve(swm_slot *s, uint64_t n) { s->reserve_n = n; }

void swm_slot_set_gate(swm_slot *s, swm_gate *g) { s->gate = g; }

/* The gate around one run's kernels: wait for the previous run (any slot)
 * on this slot's stream, and after the kernels publish this slot's end. */
struct GateTurn {
    swm_slot *sl;
    std::unique_lock<std::mutex> lk;
    explicit GateTurn(swm_slot *s) : sl(s)
    {
        if (!sl->gate) return;
        lk = std::unique_lock<std::mutex>(sl->gate->mu);
        if (sl->gate->last) (void)hipStreamWaitEvent(sl->stream, sl->gate->last, 0);
    }
    void done()
    {
        if (!sl->gate || !lk.owns_lock()) return;
        if (hipEventRecord(sl->ev_gate, sl->stream) == hipSuccess) sl->gate->last = sl->ev_gate;
        lk.unlock();
    }
    ~GateTurn() { done(); }
};

static double host_ms()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

/* The final wait of a run, with the host time split around it. */
static int final_sync(swm_slot *sl)
{
    const double t0 = host_ms();
    const hipError_t e = hipStreamSynchronize(sl->stream);
    const double t1 = host_ms();
    sl->stats.ms_host_enqueue = t0 - sl->t_entry;
    sl->stats.ms_host_wait = t1 - t0;
    return e == hipSuccess ? SWM_OK : SWM_EDEVICE;
}

/* Device pointers of one run. */
struct Pipe {
    uint64_t n = 0, nrs = 0;
    uint32_t seln = 0;
    SeedOut so;
    PlanIO pio;
    CandSet cs;
    uint64_t *surv_off = nullptr;
    uint32_t *bd_off = nullptr, *cand_off = nullptr, *trace = nullptr, *heavy_c = nullptr, *heavy_e = nullptr;
    EmitPart *epart = nullptr;
    uint32_t *rpart = nullptr;
    unsigned long long *counters = nullptr;
    uint32_t *u32c = nullptr;
    uint64_t *scan_tmp = nullptr;
    uint32_t *lbins = nullptr;       /* [LBIN_ROWS + 1][256]: candidates per read length, then each length's next place */
    uint32_t *perm = nullptr;        /* [cap]: verification order (CandSet::perm) */
};

/* counters: u64 [0] candidates, [1] records, [2] band steps, [3] surviving
 * positions; u32 view of [4..9]: [0] large batches, [1] tracebacks,
 * [2] rejected windows, [3] heavy plan items, [4] their queue head,
 * [5] heavy compaction items, [6] verification order is the identity, [7] overflow, [8] heavy plan
 * batches, [9] giant plan items, [10] giant / [11] other heavy emission
 * reads, [12] items with an irregular bucket, [13] heavy batches for k_merge */
template <class Alloc>
static int pipe_alloc_a(Alloc *ws, uint64_t n, uint32_t seln, Pipe &P)
{
    const uint64_t nrs = 2 * n;
    P.n = n;
    P.nrs = nrs;
    P.seln = seln;
    int rc;
    if ((rc = ws->get(swm_ws::WORDS, &P.so.words, nrs * 32)) || (rc = ws->get(swm_ws::SEL, &P.so.sel, nrs * seln)) ||
        (rc = ws->get(swm_ws::SELC, &P.so.selc, nrs * seln)) ||
        (rc = ws->get(swm_ws::NSEL, &P.so.nsel, nrs)) || (rc = ws->get(swm_ws::NPOS, &P.so.npos, nrs + 1)) ||
        (rc = ws->get(swm_ws::NCHUNK, &P.so.nchunk, nrs + 1)) || (rc = ws->get(swm_ws::COUNTERS, &P.counters, 12)) ||
        (rc = ws->get(swm_ws::SURV_OFF, &P.surv_off, nrs + 1)) || (rc = ws->get(swm_ws::BD_OFF, &P.bd_off, nrs + 1)) ||
        (rc = ws->get(swm_ws::NBATCH, &P.pio.nbatch, nrs)) || (rc = ws->get(swm_ws::NUNIQ, &P.pio.nuniq, nrs + 1)) ||
        (rc = ws->get(swm_ws::HEAVY, &P.pio.heavy, nrs)) || (rc = ws->get(swm_ws::GIANT, &P.pio.giant, nrs)) ||
        (rc = ws->get(swm_ws::IRREG, &P.pio.irreg, nrs)) || (rc = ws->get(swm_ws::CAND_OFF, &P.cand_off, nrs + 1)) ||
        (rc = ws->get(swm_ws::HEAVY_C, &P.heavy_c, nrs)) || (rc = ws->get(swm_ws::HEAVY_E, &P.heavy_e, 2 * n)) ||
        (rc = ws->get(swm_ws::EPART, &P.epart, 2 * n)) ||
        (rc = ws->get(swm_ws::RPART, &P.rpart, (uint64_t)blocks_for(n, 128))) ||
        (rc = ws->get(swm_ws::LBINS, &P.lbins, (LBIN_ROWS + 1) * 256u)) ||
        (rc = ws->get(swm_ws::SCAN_TMP, &P.scan_tmp, 2 * (uint64_t)blocks_for(nrs + 1, SCAN_TILE))))
        return rc;
    P.so.counters = P.counters;
    P.so.lbins = P.lbins;
    P.so.nuniq = P.pio.nuniq;
    P.so.nbatch = P.pio.nbatch;
    P.u32c = reinterpret_cast<uint32_t *>(P.counters + 4);
    P.pio.sel = P.so.sel;
    P.pio.selc = P.so.selc;
    P.pio.seln = seln;
    P.pio.nsel = P.so.nsel;
    P.pio.surv_off = P.surv_off;
    P.pio.bd_off = P.bd_off;
    P.pio.big_count = P.u32c;
    P.pio.heavy_count = P.u32c + 3;
    P.pio.heavy_head = P.u32c + 4;
    P.pio.hb_count = P.u32c + 8;
    P.pio.giant_count = P.u32c + 9;
    P.pio.irreg_count = P.u32c + 12;
    P.pio.merge_count = P.u32c + 13;
    P.pio.light_pos = P.counters + 11;
    P.pio.ovf = P.u32c + 7;
    P.cs.ovf = P.u32c + 7;
    P.cs.perm_id = P.u32c + 6;                    /* 8-byte aligned, the overflow flag next */
    if ((reinterpret_cast<uintptr_t>(P.cs.perm_id) & 7u) || P.cs.perm_id + 1 != P.cs.ovf) return SWM_EINVAL;
    return SWM_OK;
}

template <class Alloc>
static int pipe_alloc_b(Alloc *ws, uint64_t cap_pos, uint64_t cap_chunk, Pipe &P)
{
    int rc;
    if ((rc = ws->get(swm_ws::SURV, &P.pio.surv, cap_pos + 1)) ||
        (rc = ws->get(swm_ws::BDESC, &P.pio.bdesc, cap_chunk + 1)) ||
        (rc = ws->get(swm_ws::BIG, &P.pio.big, cap_chunk + 1)) ||
        (rc = ws->get(swm_ws::CDESC, &P.pio.cdesc, cap_chunk + 1)) ||
        (rc = ws->get(swm_ws::HBATCH, &P.pio.hbatch, cap_chunk + 1)) ||
        (rc = ws->get(swm_ws::MERGE, &P.pio.merge, cap_chunk + 1)))
        return rc;
    return SWM_OK;
}

template <class Alloc>
static int pipe_alloc_c(Alloc *ws, uint64_t cap, Pipe &P)
{
    int rc;
    if ((rc = ws->get(swm_ws::C_VAL, &P.cs.val, cap + 1)) || (rc = ws->get(swm_ws::C_RS, &P.cs.rs, cap + 1)) ||
        (rc = ws->get(swm_ws::C_BI, &P.cs.bi, cap + 1)) || (rc = ws->get(swm_ws::C_ERR, &P.cs.err, cap + 1)) ||
        (rc = ws->get(swm_ws::C_LOC, &P.cs.loc, cap + 1)) || (rc = ws->get(swm_ws::C_REC, &P.cs.rec, cap + 1)) ||
        (rc = ws->get(swm_ws::C_INFO, &P.cs.info, cap + 1)) || (rc = ws->get(swm_ws::TRACE, &P.trace, cap + 1)) ||
        (rc = ws->get(swm_ws::PERM, &P.perm, cap + 1)))
        return rc;
    P.cs.perm = P.perm;
    return SWM_OK;
}

static void scan_launch(ScanJob j, uint32_t n_arrays, hipStream_t st)
{
    j.nb = blocks_for(j.n, SCAN_TILE);
    k_scan_reduce<<<dim3(j.nb, n_arrays), SCAN_NT, 0, st>>>(j);
    k_scan_down<<<dim3(j.nb, n_arrays), SCAN_NT, 0, st>>>(j);
}

/* Stage A: seeds (which also zero the counters), offsets of the positions
 * and chunks with their capacity check. */
static int enq_a(const swm_ctx_t *ctx, Pipe &P, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                 uint32_t e, uint64_t cap_pos, uint64_t cap_chunk, hipStream_t st, hipEvent_t *ev)
{
    if (ev) SWM_HIP_CHECK(hipEventRecord(ev[0], st));
    P.so.irreg = ctx->d_irreg;
    P.so.n_irreg = ctx->irr_flags ? 0u : ctx->n_irreg;
    P.so.off_mask = ctx->irr_flags ? 0x7fffffffu : 0xffffffffu;
    if (ctx->k == 12 && !ctx->opt.seed_generic)
        k_seed_fixed<12><<<blocks_for(P.n, SEED_READS), 2 * SEED_READS, 0, st>>>(
            d_reads, stride, d_lens, (uint32_t)P.n, e, P.seln, ctx->d_seed_tab, ctx->d_recipe, P.so);
    else
        k_seed<<<blocks_for(P.n, SEED_READS), 2 * SEED_READS, 0, st>>>(d_reads, stride, d_lens, (uint32_t)P.n, ctx->k,
                                                                       e, P.seln, ctx->d_seed_tab, ctx->d_recipe, P.so);
    SWM_HIP_CHECK(hipGetLastError());
    if (ev) SWM_HIP_CHECK(hipEventRecord(ev[1], st));
    ScanJob j{};
    j.in0 = P.so.npos; j.out0 = P.surv_off;
    j.in1 = P.so.nchunk; j.out1 = P.bd_off;
    j.n = P.nrs + 1;
    j.sums = P.scan_tmp;
    j.first = 0;
    j.cap[0] = cap_pos; j.cap[1] = cap_chunk;
    j.tot[0] = P.counters + 3; j.tot[1] = nullptr;
    j.ovf = P.u32c + 7;
    if (ctx->opt.work_queue || ctx->n_irreg) {
        j.cls_nsel = P.so.nsel;
        j.cls_n = (uint32_t)P.nrs;
        j.cls_heavy_on = ctx->opt.work_queue;
        j.cls_heavy = P.pio.heavy; j.cls_heavy_count = P.pio.heavy_count;
        j.cls_giant = P.pio.giant; j.cls_giant_count = P.pio.giant_count;
        j.cls_irreg = P.pio.irreg; j.cls_irreg_count = P.pio.irreg_count;
        j.giant_pos = ctx->opt.giant_pos;
    }
    scan_launch(j, 2, st);
    SWM_HIP_CHECK(hipGetLastError());
    return SWM_OK;
}

/* Stage B: batch plans, large-batch sort, candidate offsets with their
 * capacity check.  In the work-queue form with plan_overlap the heavy and
 * irregular items' kernels (latency-bound chains) run on the slot's second
 * stream while the light-item kernel runs on the first: they take disjoint
 * items and meet only in the large-batch list's atomic counter. */
static int enq_b(const swm_ctx_t *ctx, swm_slot *sl, Pipe &P, uint64_t cap_cand, hipStream_t st, hipEvent_t *ev)
{
    const swm_opts &o = ctx->opt;
    if (o.work_queue) {
        const bool fork = o.plan_overlap != 0;
        /* heavy items: seed-heap replay, batch fetch; irregular items and
         * the merges of descending chunks */
        auto heavy_chain = [&](hipStream_t hs) {
            k_plan_heavy<<<ctx->n_cu * o.plan_heavy_waves, 64, 0, hs>>>(ctx->d_pl, P.pio);
            k_plan_fetch<<<ctx->n_cu * 32, 64, 0, hs>>>(ctx->d_pl, P.pio);
            if (ctx->n_irreg) {
                k_plan_irreg<<<ctx->n_cu * 4, 64, 0, hs>>>(ctx->d_pl, P.pio);
                k_merge<<<ctx->n_cu * 4, 64, 0, hs>>>(P.pio);
            }
        };
        if (fork) {
            SWM_HIP_CHECK(hipEventRecord(sl->ev_fork, st));
            SWM_HIP_CHECK(hipStreamWaitEvent(sl->side, sl->ev_fork, 0));
            heavy_chain(sl->side);
            SWM_HIP_CHECK(hipEventRecord(sl->ev_join, sl->side));
        }
        if (ev) SWM_HIP_CHECK(hipEventRecord(ev[6], st));
        const uint32_t pgrid = (uint32_t)std::min<uint64_t>(P.n ? P.n : 1, (uint64_t)ctx->n_cu * o.plan_light_waves);
        P.pio.heavy_cands = o.heavy_cands;      /* as k_compact gets it in the work-queue form */
        if (P.seln <= 32)       /* seln <= 64 for every e */
            k_plan_pair<32><<<pgrid, 64, 0, st>>>((uint32_t)P.n, ctx->d_pl, P.pio);
        else
            k_plan_pair<64><<<pgrid, 64, 0, st>>>((uint32_t)P.n, ctx->d_pl, P.pio);
        if (ev) SWM_HIP_CHECK(hipEventRecord(ev[7], st));
        if (fork) SWM_HIP_CHECK(hipStreamWaitEvent(st, sl->ev_join, 0));
        else heavy_chain(st);
    } else {
        k_plan_plain<<<blocks_for(P.nrs, 64), 64, 0, st>>>((uint32_t)P.nrs, ctx->d_pl, P.pio);
        if (ctx->n_irreg) k_plan_irreg<<<ctx->n_cu * 4, 64, 0, st>>>(ctx->d_pl, P.pio);
    }
    k_bigsort<<<ctx->n_cu * 32, 64, 0, st>>>(P.pio.big, P.pio.big_count, P.surv_off, P.bd_off, P.pio.surv, P.pio.bdesc,
                                    P.pio.nuniq);
    SWM_HIP_CHECK(hipGetLastError());
    if (ev) SWM_HIP_CHECK(hipEventRecord(ev[2], st));
    ScanJob j{};
    j.in1 = P.pio.nuniq; j.out1 = P.cand_off;
    j.n = P.nrs + 1;
    j.sums = P.scan_tmp;
    j.first = 1;
    j.cap[0] = ~0ull; j.cap[1] = cap_cand;
    j.tot[0] = nullptr; j.tot[1] = P.counters;
    j.ovf = P.u32c + 7;
    scan_launch(j, 1, st);
    SWM_HIP_CHECK(hipGetLastError());
    return SWM_OK;
}

/* Stage C: candidate list, verification, records, per-read results.
 * grid_cand sizes the candidate kernels (their count is read on the device). */
static int enq_c(const swm_ctx_t *ctx, Pipe &P, const uint32_t *d_lens, uint32_t e, swm_read_result_t *d_out,
                 uint64_t grid_cand, hipStream_t st, hipEvent_t *ev)
{
    const swm_opts &o = ctx->opt;
    const uint32_t heavy_cands = o.work_queue ? o.heavy_cands : 0xffffffffu;
    const uint32_t *ncand = P.cand_off + P.nrs;
    k_compact<<<blocks_for(P.nrs, 256), 256, 0, st>>>((uint32_t)P.nrs, P.surv_off, P.pio.surv, P.bd_off, P.pio.bdesc,
                                                      P.pio.nbatch, P.cand_off, P.cs, heavy_cands, P.heavy_c,
                                                      P.u32c + 5);
    if (o.work_queue)
        k_compact_heavy<<<ctx->n_cu * 8, 256, 0, st>>>(P.heavy_c, P.u32c + 5, P.surv_off, P.pio.surv, P.bd_off,
                                                       P.pio.bdesc, P.pio.nbatch, P.cand_off, P.cs);
    if (P.cs.perm) {
        /* verification order by read length (the identity for one length) */
        const uint32_t lg = (uint32_t)ctx->n_cu * LSORT_WG_PER_CU;
        k_lcount<<<lg, 256, 0, st>>>((uint32_t)P.nrs, P.cand_off, d_lens, P.cs.ovf, P.lbins);
        k_lscan<<<1, 256, 0, st>>>(P.cs.ovf, P.lbins, P.u32c + 6);
        k_lscatter<<<lg, 256, 0, st>>>((uint32_t)P.nrs, P.cand_off, d_lens, P.cs.ovf, P.u32c + 6, P.lbins, P.perm);
    }
    SWM_HIP_CHECK(hipGetLastError());
    if (ev) SWM_HIP_CHECK(hipEventRecord(ev[3], st));
    const uint32_t mb = o.myers_block;
    if (o.myers_packed && e <= SWM_MAX_E_PACKED)
        k_myers2<<<blocks_for((grid_cand ? grid_cand : 1) / 2 + 1, mb), mb, 0, st>>>(ncand, d_lens, P.so.words,
                                                                                    ctx->d_base_2bits,
                                                                                    ctx->ref_length, e, P.cs, P.trace,
                                                                                    P.u32c + 1, P.counters + 2,
                                                                                    P.u32c + 2);
    else
        k_myers<<<blocks_for(grid_cand ? grid_cand : 1, mb), mb, 0, st>>>(ncand, d_lens, P.so.words,
                                                                          ctx->d_base_2bits, ctx->ref_length, e,
                                                                          P.cs, P.trace, P.u32c + 1, P.counters + 2,
                                                                          P.u32c + 2);
    SWM_HIP_CHECK(hipGetLastError());
    if (ev) SWM_HIP_CHECK(hipEventRecord(ev[4], st));
    /* persistent: 8 waves per SIMD, the traceback chains are latency-bound */
    (P.cs.runs ? k_trace<true> : k_trace<false>)<<<ctx->n_cu * 32, 64, 0, st>>>(P.trace, P.u32c + 1, d_lens, P.so.words, ctx->d_base_2bits, ctx->ref_length, e,
                                          P.cs);
    k_emit<<<blocks_for(P.n, 128), 128, 0, st>>>((uint32_t)P.n, e, P.cand_off, P.cs, d_out, P.rpart,
                                                 heavy_cands, o.giant_cands, P.heavy_e, P.u32c + 10);
    if (o.work_queue)
        (P.cs.wr ? k_emit_heavy<true> : k_emit_heavy<false>)<<<ctx->n_cu * o.emit_heavy_waves, 64, 0, st>>>(
            P.heavy_e, P.u32c + 10, (uint32_t)P.n, e, P.cand_off, P.cs, P.epart, P.counters + 1);
    k_emit_combine<<<ctx->n_cu * 2, 256, 0, st>>>(P.heavy_e, P.u32c + 10, (uint32_t)P.n, P.epart, d_out, P.rpart,
                                                  blocks_for(P.n, 128), P.counters + 1);
    SWM_HIP_CHECK(hipGetLastError());
    if (ev) SWM_HIP_CHECK(hipEventRecord(ev[5], st));
    return SWM_OK;
}

static void stats_from_counters(swm_stats_t *s, uint64_t n, const unsigned long long *hc)
{
    const uint32_t *u = reinterpret_cast<const uint32_t *>(hc + 4);
    s->n_reads = n;
    s->n_verifications = hc[0] - u[2];
    s->n_band_steps = hc[2];
    s->n_records = hc[1];
    s->n_positions = hc[3];
    const uint32_t *u2 = reinterpret_cast<const uint32_t *>(hc + 8);   /* u32 view of [8..9] */
    s->n_large_batches = u[0];
    s->n_heavy_items = u[3] + u2[1];
    s->n_heavy_reads = u2[2] + u2[3];
    s->n_giant_items = u2[1];
    s->n_giant_reads = u2[2];
    s->n_irregular_items = reinterpret_cast<const uint32_t *>(hc + 10)[0];
    s->n_light_positions = hc[11];
}

/* The per-read candidate histogram of the run just finished (stats). */
static int collect_hist(const swm_ctx_t *ctx, swm_slot *sl, const Pipe &P)
{
    unsigned long long *d_hist;
    int rc;
    if ((rc = sl->ws.get(swm_ws::HIST, &d_hist, 32))) return rc;
    hipStream_t st = sl->stream;
    SWM_HIP_CHECK(hipMemsetAsync(d_hist, 0, 32 * sizeof(unsigned long long), st));
    k_hist<<<std::min<uint32_t>(blocks_for(P.n, 256), (uint32_t)ctx->n_cu * 4), 256, 0, st>>>((uint32_t)P.n, P.cand_off,
                                                                                             d_hist);
    SWM_HIP_CHECK(hipGetLastError());
    unsigned long long h[25];
    SWM_HIP_CHECK(hipMemcpyAsync(h, d_hist, sizeof(h), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    for (int b = 0; b < 24; b++) sl->stats.cand_hist[b] = h[b];
    sl->stats.max_cands = h[24];
    return SWM_OK;
}

static void observe(swm_slot *sl, uint64_t n, uint64_t tot_pos, uint64_t tot_chunk, uint64_t ncand)
{
    if (!n) return;
    const double dn = (double)n;
    if (tot_pos / dn > sl->r_pos) sl->r_pos = tot_pos / dn;
    if (tot_chunk / dn > sl->r_chunk) sl->r_chunk = tot_chunk / dn;
    if (ncand / dn > sl->r_cand) sl->r_cand = ncand / dn;
}

/* One launch per stage, every buffer sized from the totals read back after
 * stages A and B (two host round trips). */
/* h_out (page-locked, may be NULL): the first n_out results are copied there
 * before the run's last synchronisation. */
static int records_out(const swm_ctx_t *ctx, swm_slot *sl, Pipe &P, const uint32_t *d_lens, uint32_t ncand,
                       swm_records_t *recs);

/* recs (may be NULL): also collect every record and its CIGAR text. */
static int run_synced(swm_ctx_t *ctx, swm_slot *sl, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                      uint64_t n, uint32_t e, swm_read_result_t *d_out, swm_read_result_t *h_out, uint64_t n_out,
                      swm_records_t *recs = nullptr)
{
    hipStream_t st = sl->stream;
    swm_ws *ws = &sl->ws;
    memset(&sl->stats, 0, sizeof(sl->stats));
    sl->stats.n_reads = n;
    if (n == 0) return SWM_OK;
    const uint32_t seln = 4 * ((e + 1) < 16 ? (e + 1) : 16);
    Pipe P;
    int rc;
    GateTurn turn(sl);
    if ((rc = pipe_alloc_a(ws, n, seln, P))) return rc;
    if ((rc = enq_a(ctx, P, d_reads, stride, d_lens, e, ~0ull, ~0ull, st, sl->ev))) return rc;
    uint64_t *hs = sl->h_small;
    hs[9] = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(hs + 8, P.surv_off + P.nrs, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipMemcpyAsync(hs + 9, P.bd_off + P.nrs, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    const uint64_t tot_pos = hs[8], tot_chunk = (uint32_t)hs[9];
    if ((rc = pipe_alloc_b(ws, tot_pos, tot_chunk, P))) return rc;
    if ((rc = enq_b(ctx, sl, P, ~0ull, st, sl->ev))) return rc;
    hs[10] = 0;
    SWM_HIP_CHECK(hipMemcpyAsync(hs + 10, P.cand_off + P.nrs, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    const uint32_t ncand = (uint32_t)hs[10];
    if ((rc = pipe_alloc_c(ws, ncand, P))) return rc;
    if (recs && ((rc = ws->get(swm_ws::R_WR, &P.cs.wr, (uint64_t)ncand + 1)) ||
                 (rc = ws->get(swm_ws::R_TB, &P.cs.tb, (uint64_t)ncand + 1)) ||
                 (rc = ws->get(swm_ws::R_RUNS, &P.cs.runs, ((uint64_t)ncand + 1) * RUN_SLOT))))
        return rc;
    if ((rc = enq_c(ctx, P, d_lens, e, d_out, ncand, st, sl->ev))) return rc;
    if (recs && (rc = records_out(ctx, sl, P, d_lens, ncand, recs))) return rc;
    turn.done();
    SWM_HIP_CHECK(hipMemcpyAsync(hs, P.counters, 12 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (h_out && n_out)
        SWM_HIP_CHECK(hipMemcpyAsync(h_out, d_out, n_out * sizeof(swm_read_result_t), hipMemcpyDeviceToHost, st));
    if ((rc = final_sync(sl))) return rc;
    stats_from_counters(&sl->stats, n, reinterpret_cast<const unsigned long long *>(hs));
    observe(sl, n, tot_pos, tot_chunk, ncand);
    if (ctx->opt.histogram && (rc = collect_hist(ctx, sl, P)) != SWM_OK) return rc;
    float ms;
    (void)hipEventElapsedTime(&ms, sl->ev[0], sl->ev[1]); sl->stats.ms_seed = ms;
    (void)hipEventElapsedTime(&ms, sl->ev[1], sl->ev[2]); sl->stats.ms_plan = ms;
    (void)hipEventElapsedTime(&ms, sl->ev[2], sl->ev[3]); sl->stats.ms_compact = ms;
    (void)hipEventElapsedTime(&ms, sl->ev[3], sl->ev[4]); sl->stats.ms_myers = ms;
    (void)hipEventElapsedTime(&ms, sl->ev[4], sl->ev[5]); sl->stats.ms_emit = ms;
    (void)hipEventElapsedTime(&ms, sl->ev[0], sl->ev[5]); sl->stats.ms_total = ms;
    sl->stats.ms_plan_light = 0;
    if (ctx->opt.work_queue) { (void)hipEventElapsedTime(&ms, sl->ev[6], sl->ev[7]); sl->stats.ms_plan_light = ms; }
    return SWM_OK;
}

/* The records of the run just enqueued: ranks and text offsets by one scan
 * over the candidates, the totals read back, then k_records and the copies
 * to library-owned host memory. */
static int records_out(const swm_ctx_t *ctx, swm_slot *sl, Pipe &P, const uint32_t *d_lens, uint32_t ncand,
                       swm_records_t *recs)
{
    memset(recs, 0, sizeof(*recs));
    if (ncand == 0) return SWM_OK;
    swm_ws *ws = &sl->ws;
    hipStream_t st = sl->stream;
    uint64_t *ridx, *sums;
    uint32_t *toff;
    unsigned long long *tot;
    int rc;
    const uint32_t nb = blocks_for(ncand, SCAN_TILE);
    if ((rc = ws->get(swm_ws::R_IDX, &ridx, (uint64_t)ncand + 1)) || (rc = ws->get(swm_ws::R_TOFF, &toff, (uint64_t)ncand + 1)) ||
        (rc = ws->get(swm_ws::R_SCAN, &sums, 2ull * nb)) || (rc = ws->get(swm_ws::R_TOT, &tot, 4)))
        return rc;
    ScanJob j{};
    j.in0 = P.cs.wr; j.out0 = ridx;
    j.in1 = P.cs.tb; j.out1 = toff;
    j.n = ncand;
    j.sums = sums;
    j.first = 0;
    j.cap[0] = ~0ull; j.cap[1] = 0xffffffffull;
    j.tot[0] = tot; j.tot[1] = tot + 1;
    j.ovf = reinterpret_cast<uint32_t *>(tot + 2);
    SWM_HIP_CHECK(hipMemsetAsync(tot, 0, 4 * sizeof(unsigned long long), st));
    scan_launch(j, 2, st);
    SWM_HIP_CHECK(hipGetLastError());
    uint64_t *hs = sl->h_small;
    SWM_HIP_CHECK(hipMemcpyAsync(hs + 11, tot, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    SWM_HIP_CHECK(hipStreamSynchronize(st));
    const uint64_t nrec = (uint32_t)hs[11], nblk = hs[11] >> 32, nbytes = hs[12];
    if ((uint32_t)hs[13]) return SWM_ENOMEM;           /* CIGAR text beyond 4 GiB in one batch */
    if (nrec == 0) return SWM_OK;
    swm_record_t *d_rec;
    char *d_text;
    uint32_t *d_blk;
    if ((rc = ws->get(swm_ws::R_OUT, &d_rec, nrec)) || (rc = ws->get(swm_ws::R_TEXT, &d_text, nbytes + 1)) ||
        (rc = ws->get(swm_ws::R_BLOCK, &d_blk, nblk)))
        return rc;
    const uint32_t grid = std::min<uint32_t>(blocks_for(ncand, 256), (uint32_t)ctx->n_cu * 16);
    k_records<<<grid, 256, 0, st>>>(P.cand_off + P.nrs, d_lens, P.cand_off, P.cs, ridx, toff, d_rec, d_text, d_blk);
    SWM_HIP_CHECK(hipGetLastError());
    recs->rec = (swm_record_t *)malloc(nrec * sizeof(swm_record_t));
    recs->cigar = (char *)malloc(nbytes + 1);
    recs->block = (uint32_t *)malloc(nblk * sizeof(uint32_t));
    if (!recs->rec || !recs->cigar || !recs->block) {
        swm_records_free(recs);
        (void)hipStreamSynchronize(st);
        return SWM_ENOMEM;
    }
    if (hipMemcpyAsync(recs->rec, d_rec, nrec * sizeof(swm_record_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(recs->cigar, d_text, nbytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(recs->block, d_blk, nblk * sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        swm_records_free(recs);
        return SWM_EDEVICE;
    }
    recs->cigar[nbytes] = 0;
    recs->n = nrec;
    recs->text_bytes = nbytes;
    recs->n_blocks = nblk;
    return SWM_OK;
}

static bool same_opts(const swm_opts &a, const swm_opts &b) { return memcmp(&a, &b, sizeof(swm_opts)) == 0; }

/* Fixed-capacity form: the whole lookup -> filter -> verify -> emit pipeline
 * captured once per (buffers, n, e) into a hipGraph and replayed with no host
 * round trip.  Capacities come from the largest per-read totals seen so far
 * with 25% headroom; a batch that would overflow them stands down on the
 * device and is run again in the synced form, which also raises them. */
static int run_graph(swm_ctx_t *ctx, swm_slot *sl, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                     uint64_t n, uint32_t e, swm_read_result_t *d_out, swm_read_result_t *h_out, uint64_t n_out)
{
    if (n == 0 || sl->r_cand == 0) return run_synced(ctx, sl, d_reads, stride, d_lens, n, e, d_out, h_out, n_out);
    hipStream_t st = sl->stream;
    swm_slot::GraphEnt *ent = nullptr;
    for (auto &x : sl->g)
        if (x.used && x.reads == d_reads && x.lens == d_lens && x.out == d_out && x.n == n && x.stride == stride &&
            x.e == e && same_opts(x.opt, ctx->opt))
            ent = &x;
    const uint32_t seln = 4 * ((e + 1) < 16 ? (e + 1) : 16);
    Pipe P;
    int rc;
    bool captured = false;
    if (!ent || ent->gen != sl->ws.gen) {
        if (!ent) {
            ent = &sl->g[0];
            for (auto &x : sl->g) {
                if (!x.used) { ent = &x; break; }
                if (x.tick < ent->tick) ent = &x;
            }
        }
        if (ent->exec) { (void)hipGraphExecDestroy(ent->exec); ent->exec = nullptr; }
        ent->used = false;
        auto caps = [&](uint64_t m, uint64_t *cp, uint64_t *cc, uint64_t *cd) {
            const double dm = (double)m;
            *cp = (uint64_t)(dm * sl->r_pos * 1.25) + 4096;
            *cc = (uint64_t)(dm * sl->r_chunk * 1.25) + 1024;
            *cd = (uint64_t)(dm * sl->r_cand * 1.25) + 4096;
        };
        /* capacities: the buffers' (sized for the reserve when there is one),
         * which bound the overflow checks; the candidate grid follows n (its
         * loops are grid-stride) */
        const uint64_t n_cap = sl->reserve_n > n ? sl->reserve_n : n;
        uint64_t cap_pos, cap_chunk, cap_cand, g_pos, g_chunk, g_cand;
        caps(n_cap, &cap_pos, &cap_chunk, &cap_cand);
        caps(n, &g_pos, &g_chunk, &g_cand);
        if (n_cap > n && (rc = pipe_alloc_a(&sl->ws, n_cap, seln, P))) return rc;
        if ((rc = pipe_alloc_a(&sl->ws, n, seln, P)) || (rc = pipe_alloc_b(&sl->ws, cap_pos, cap_chunk, P)) ||
            (rc = pipe_alloc_c(&sl->ws, cap_cand, P)))
            return rc;
        hipGraph_t graph = nullptr;
        hipError_t ie;
        {
            /* a capture of a server's slot holds the slots' gate: no other
             * slot enqueues its run while a graph is captured and
             * instantiated */
            std::unique_lock<std::mutex> glk;
            if (sl->gate) glk = std::unique_lock<std::mutex>(sl->gate->mu);
            SWM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int crc = enq_a(ctx, P, d_reads, stride, d_lens, e, cap_pos, cap_chunk, st, nullptr);
            if (!crc) crc = enq_b(ctx, sl, P, cap_cand, st, nullptr);
            if (!crc) crc = enq_c(ctx, P, d_lens, e, d_out, g_cand, st, nullptr);
            const hipError_t ce = hipStreamEndCapture(st, &graph);
            if (crc || ce != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                return crc ? crc : SWM_EDEVICE;
            }
            ie = hipGraphInstantiate(&ent->exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
        }
        if (ie != hipSuccess) { ent->exec = nullptr; return SWM_EDEVICE; }
        captured = true;
        ent->used = true;
        ent->reads = d_reads; ent->lens = d_lens; ent->out = d_out;
        ent->n = n; ent->stride = stride; ent->e = e; ent->opt = ctx->opt; ent->gen = sl->ws.gen;
    } else if ((rc = pipe_alloc_a(&sl->ws, n, seln, P))) {
        return rc;   /* no growth here (the generation matched): only the counters' address is needed */
    }
    ent->tick = ++sl->tick;
    memset(&sl->stats, 0, sizeof(sl->stats));
    {
        GateTurn turn(sl);
        SWM_HIP_CHECK(hipEventRecord(sl->ev[0], st));
        SWM_HIP_CHECK(hipGraphLaunch(ent->exec, st));
        SWM_HIP_CHECK(hipEventRecord(sl->ev[5], st));
    }
    const unsigned long long *hc = reinterpret_cast<const unsigned long long *>(sl->h_small);
    SWM_HIP_CHECK(hipMemcpyAsync(sl->h_small, P.counters, 12 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (h_out && n_out)
        SWM_HIP_CHECK(hipMemcpyAsync(h_out, d_out, n_out * sizeof(swm_read_result_t), hipMemcpyDeviceToHost, st));
    if ((rc = final_sync(sl))) return rc;
    if (reinterpret_cast<const uint32_t *>(hc + 4)[7]) {
        /* over capacity: nothing was written; the synced form maps the batch
         * and records the larger totals, and the graph is captured again */
        (void)hipGraphExecDestroy(ent->exec);
        ent->exec = nullptr;
        ent->used = false;
        return run_synced(ctx, sl, d_reads, stride, d_lens, n, e, d_out, h_out, n_out);
    }
    stats_from_counters(&sl->stats, n, hc);
    sl->stats.n_graph_captures = captured ? 1 : 0;
    float ms;
    (void)hipEventElapsedTime(&ms, sl->ev[0], sl->ev[5]);
    sl->stats.ms_total = ms;
    return SWM_OK;
}

/* ------------------------------------------ no-allocation form (scratch) */

/* The device scratch of swm_map_batch_device_noalloc: one array per device
 * (the code object's own), held until the process ends.  Host code never
 * passes the array's name to a kernel (on the host it is only a stand-in);
 * each device's address comes from hipGetSymbolAddress, once per device,
 * at context creation. */
static __device__ __align__(256) uint8_t scratch[SWM_NOALLOC_SCRATCH_BYTES];

#define SCRATCH_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[SCRATCH_MAX_DEVICES];

/* The scratch address on device dev, which must be current. */
static int scratch_address(int dev, uint8_t **out)
{
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return SWM_EDEVICE;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SWM_EDEVICE;
        g_scratch_ptr[dev] = (uint8_t *)p;
    }
    *out = g_scratch_ptr[dev];
    return SWM_OK;
}

/* Bump allocator over the scratch: every buffer 256-byte aligned, nothing
 * freed (a run carves its buffers afresh). */
struct ScratchArena {
    uint8_t *base;
    size_t size, used;
    template <typename T>
    int get(int, T **out, size_t n)
    {
        const size_t bytes = (n ? n : 1) * sizeof(T);
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > size || bytes > size - off) return SWM_ENOSPC;
        *out = reinterpret_cast<T *>(base + off);
        used = off + bytes;
        return SWM_OK;
    }
};

/* What is left after the per-read arrays goes to the capacities: 60% to
 * bucket chunks (60 bytes each: descriptors, large-batch, heavy-chunk and
 * merge lists, heavy batch records), 25% to positions (4 bytes), 15% to
 * candidates (36 bytes: nine per-candidate arrays).  A batch over any of
 * them stands down on the device (the fixed-capacity checks of the graph
 * form) and the call reports SWM_ENOSPC. */
static void scratch_caps(size_t left, uint64_t *cap_pos, uint64_t *cap_chunk, uint64_t *cap_cand)
{
    const size_t slack = 16 * 256 + 256;   /* alignment of 15 buffers and their one extra element */
    const uint64_t r = left > slack ? left - slack : 0;
    *cap_chunk = r * 6 / 10 / 60;
    *cap_pos = r / 4 / 4;
    *cap_cand = r * 15 / 100 / 36;
}

/* The fixed-capacity pipeline (the graph form's kernels, enqueued directly
 * on the context's stream) over buffers carved from the scratch.  Nothing
 * is allocated; the counters come back through the slot's page-locked
 * staging, allocated at context creation.  Every buffer that must start at
 * zero is cleared by the pipeline's own kernels (k_seed zeroes the
 * counters), as in every graph replay. */
static int run_noalloc(swm_ctx_t *ctx, const char *d_reads, uint64_t stride, const uint32_t *d_lens, uint64_t n,
                       uint32_t e, swm_read_result_t *d_out)
{
    swm_slot *sl = ctx->slot0;
    hipStream_t st = sl->stream;
    memset(&sl->stats, 0, sizeof(sl->stats));
    sl->stats.n_reads = n;
    if (n == 0) return SWM_OK;
    sl->t_entry = host_ms();
    const uint32_t seln = 4 * ((e + 1) < 16 ? (e + 1) : 16);
    ScratchArena ar{ctx->d_scratch, SWM_NOALLOC_SCRATCH_BYTES, 0};
    Pipe P;
    int rc;
    if ((rc = pipe_alloc_a(&ar, n, seln, P))) return rc;
    uint64_t cap_pos, cap_chunk, cap_cand;
    scratch_caps(ar.size - ar.used, &cap_pos, &cap_chunk, &cap_cand);
    if ((rc = pipe_alloc_b(&ar, cap_pos, cap_chunk, P)) || (rc = pipe_alloc_c(&ar, cap_cand, P))) return rc;
    {
        GateTurn turn(sl);
        SWM_HIP_CHECK(hipEventRecord(sl->ev[0], st));
        if ((rc = enq_a(ctx, P, d_reads, stride, d_lens, e, cap_pos, cap_chunk, st, nullptr)) ||
            (rc = enq_b(ctx, sl, P, cap_cand, st, nullptr)) || (rc = enq_c(ctx, P, d_lens, e, d_out, cap_cand, st, nullptr)))
            return rc;
        SWM_HIP_CHECK(hipEventRecord(sl->ev[5], st));
    }
    const unsigned long long *hc = reinterpret_cast<const unsigned long long *>(sl->h_small);
    SWM_HIP_CHECK(hipMemcpyAsync(sl->h_small, P.counters, 12 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if ((rc = final_sync(sl))) return rc;
    if (reinterpret_cast<const uint32_t *>(hc + 4)[7]) return SWM_ENOSPC;   /* over a capacity: nothing written */
    stats_from_counters(&sl->stats, n, hc);
    float ms;
    (void)hipEventElapsedTime(&ms, sl->ev[0], sl->ev[5]);
    sl->stats.ms_total = ms;
    return SWM_OK;
}

int swm_slot_run_device(swm_ctx_t *ctx, swm_slot *sl, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                        uint64_t n, uint32_t e, swm_read_result_t *d_out, bool graph)
{
    if (2 * n >= 0xfffffff0ull) return SWM_EINVAL;
    sl->t_entry = host_ms();
    return graph ? run_graph(ctx, sl, d_reads, stride, d_lens, n, e, d_out, nullptr, 0)
                 : run_synced(ctx, sl, d_reads, stride, d_lens, n, e, d_out, nullptr, 0);
}

static bool is_pinned(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   /* plain pageable memory: clear the status */
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

/* Host batches in and out through the slot's own device buffers: the
 * segments are mapped together in one pipeline run, back to back in the
 * slot's read buffer (every read maps on its own, so a read's result does
 * not depend on its neighbours).  Reads from the total up to n_pad are given
 * length 0 (they map to nothing), so runs of nearby sizes share one captured
 * graph.  Page-locked input is copied by DMA straight from the caller's
 * memory, pageable input through the slot's page-locked staging; results
 * come back through that staging. */
int swm_slot_run_segments(swm_ctx_t *ctx, swm_slot *sl, const swm_segment_t *seg, uint32_t nseg, uint64_t stride,
                          uint64_t n_pad, uint32_t e, bool graph)
{
    uint64_t n = 0;
    for (uint32_t i = 0; i < nseg; i++) n += seg[i].n;
    if (n_pad < n) n_pad = n;
    if (n == 0) { memset(&sl->stats, 0, sizeof(sl->stats)); return SWM_OK; }
    if (2 * n_pad >= 0xfffffff0ull) return SWM_EINVAL;
    sl->t_entry = host_ms();
    char *d_reads;
    uint32_t *d_lens;
    swm_read_result_t *d_out;
    int rc;
    const uint64_t n_buf = n_pad > sl->reserve_n ? n_pad : sl->reserve_n;
    if ((rc = sl->ws.get(swm_ws::IN_READS, &d_reads, n_buf * stride + 4)) ||
        (rc = sl->ws.get(swm_ws::IN_LENS, &d_lens, n_buf)) || (rc = sl->ws.get(swm_ws::OUT, &d_out, n_buf)) ||
        (rc = sl->pin(sl->h_out, n * sizeof(swm_read_result_t))))
        return rc;
    hipStream_t st = sl->stream;
    uint64_t off = 0;
    for (uint32_t i = 0; i < nseg; i++) {
        const uint64_t m = seg[i].n;
        if (!m) continue;
        const char *src_r = seg[i].reads;
        const uint32_t *src_l = seg[i].lens;
        if (!is_pinned(src_r)) {
            if ((rc = sl->pin(sl->h_reads, n * stride))) return rc;
            char *h = (char *)sl->h_reads.p + off * stride;
            memcpy(h, src_r, m * stride);
            src_r = h;
        }
        if (!is_pinned(src_l)) {
            if ((rc = sl->pin(sl->h_lens, n * 4))) return rc;
            uint32_t *h = (uint32_t *)sl->h_lens.p + off;
            memcpy(h, src_l, m * 4);
            src_l = h;
        }
        SWM_HIP_CHECK(hipMemcpyAsync(d_reads + off * stride, src_r, m * stride, hipMemcpyHostToDevice, st));
        SWM_HIP_CHECK(hipMemcpyAsync(d_lens + off, src_l, m * 4, hipMemcpyHostToDevice, st));
        off += m;
    }
    if (n_pad > n) SWM_HIP_CHECK(hipMemsetAsync(d_lens + n, 0, (n_pad - n) * 4, st));
    swm_read_result_t *h_out = (swm_read_result_t *)sl->h_out.p;
    rc = graph ? run_graph(ctx, sl, d_reads, stride, d_lens, n_pad, e, d_out, h_out, n)
               : run_synced(ctx, sl, d_reads, stride, d_lens, n_pad, e, d_out, h_out, n);
    if (rc) return rc;
    sl->stats.n_reads = n;
    off = 0;
    for (uint32_t i = 0; i < nseg; i++) {
        memcpy(seg[i].out, h_out + off, seg[i].n * sizeof(swm_read_result_t));
        off += seg[i].n;
    }
    return SWM_OK;
}

int swm_slot_run_host(swm_ctx_t *ctx, swm_slot *sl, const char *reads, uint64_t stride, const uint32_t *lens,
                      uint64_t n, uint64_t n_pad, uint32_t e, swm_read_result_t *out, bool graph)
{
    const swm_segment_t seg = {reads, lens, n, out};
    return swm_slot_run_segments(ctx, sl, &seg, 1, stride, n_pad, e, graph);
}

/* The device keeps each bucket entry as one 8-byte (position, residual)
 * pair, so a bucket chunk costs one random cache line, not two. */
/* Places of the buckets in the device array (d_order[c], in entries of 8
 * bytes): each k-mer's bucket directly before its reverse complement's (a
 * pair; a reverse-palindromic k-mer's bucket alone), every pair of at most 16
 * entries inside one 128-byte line, so that a fetch of a pair reads one line
 * where pairs laid back to back straddle two about a third of the time.
 * Best fit decreasing: pairs over 16 entries start on a line of their own,
 * then the others, largest first, go into the fullest line with room for
 * them.  Empty pairs point at entry 0.  Returns the array's length in
 * entries, or 0 when it would reach `limit` (the offsets must stay below). */
static uint64_t pair_layout(const uint32_t *hash_count, uint64_t n_buckets, uint32_t k, uint32_t *d_order,
                            uint64_t limit)
{
    std::vector<uint32_t> by_size[17];
    std::vector<uint32_t> big;
    for (uint64_t i = 0; i < n_buckets; i++) {
        const uint32_t rci = rc_bases((uint32_t)i, k);
        if (rci < i) continue;
        const uint64_t n = (uint64_t)hash_count[i] + (rci != i ? hash_count[rci] : 0u);
        if (n == 0) {
            d_order[i] = 0;
            d_order[rci] = 0;
        } else if (n > 16) {
            big.push_back((uint32_t)i);
        } else {
            by_size[n].push_back((uint32_t)i);
        }
    }
    auto place = [&](uint32_t i, uint64_t at) {
        const uint32_t rci = rc_bases(i, k);
        d_order[i] = (uint32_t)at;
        if (rci != i) d_order[rci] = (uint32_t)(at + hash_count[i]);
    };
    std::vector<uint64_t> open[16];     /* open[f]: lines with f free entries at their end */
    uint64_t lines = 0;
    for (uint32_t i : big) {
        const uint32_t rci = rc_bases(i, k);
        const uint64_t n = (uint64_t)hash_count[i] + (rci != i ? hash_count[rci] : 0u);
        if ((lines + (n + 15) / 16) * 16 >= limit) return 0;
        place(i, lines * 16);
        lines += (n + 15) / 16;
        if (n % 16) open[16 - n % 16].push_back(lines - 1);
    }
    for (uint32_t n = 16; n >= 1; n--) {
        for (uint32_t i : by_size[n]) {
            uint32_t f = n;
            while (f < 16 && open[f].empty()) f++;
            uint64_t line;
            if (f < 16) {
                line = open[f].back();
                open[f].pop_back();
            } else {
                if ((lines + 1) * 16 >= limit) return 0;
                line = lines++;
            }
            place(i, line * 16 + (16 - f));
            if (f > n) open[f - n].push_back(line);
        }
    }
    return lines * 16;
}

static int upload_pl(uint2 *d, const uint32_t *pos, const uint16_t *ls, uint64_t n,
                     const std::vector<uint64_t> &descents)
{
    const uint64_t step = 1ull << 22;
    uint2 *h = (uint2 *)malloc((n < step ? n : step) * sizeof(uint2));
    if (!h) return SWM_ENOMEM;
    int rc = SWM_OK;
    size_t di = 0;
    for (uint64_t b = 0; b < n && rc == SWM_OK; b += step) {
        uint64_t m = n - b < step ? n - b : step;
        for (uint64_t i = 0; i < m; i++) h[i] = make_uint2(pos[b + i], ls[b + i]);
        for (; di < descents.size() && descents[di] < b + m; di++) h[descents[di] - b].y |= PL_DESCENT;
        if (hipMemcpy(d + b, h, m * sizeof(uint2), hipMemcpyHostToDevice) != hipSuccess) rc = SWM_EDEVICE;
    }
    free(h);
    return rc;
}

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
    swm_ctx_t *c = new (std::nothrow) swm_ctx_t();
    if (!c) return SWM_ENOMEM;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return SWM_EDEVICE; }
    if (scratch_address(c->device, &c->d_scratch) != SWM_OK) { delete c; return SWM_EDEVICE; }
    c->slot0 = swm_slot_create(c);
    if (!c->slot0) { swm_ctx_destroy(c); return SWM_EDEVICE; }
    c->k = k;
    c->n_cu = 256;
    (void)hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device);
    c->n_buckets = 1ull << (2 * k);
    c->n_pos = n_pos;
    c->n_2bits_words = n_2bits_words;
    c->ref_length = 16ull * (n_2bits_words * 4ull);
    /* read-encoding recipes for every length (reads of length 0 map to nothing) */
    swm_recipe_t *rec = (swm_recipe_t *)calloc(SWM_RECIPE_MAX_L + 1, sizeof(swm_recipe_t));
    if (!rec) { swm_ctx_destroy(c); return SWM_ENOMEM; }
    for (uint32_t L = 0; L <= SWM_RECIPE_MAX_L; L++) {
        if (L / (k + SWM_SEED_STEP - 1) > 16) continue;       /* such reads are rejected before use */
        if (swm_recipe_build(L, k, &rec[L]) != 0) { free(rec); swm_ctx_destroy(c); return SWM_EINVAL; }
    }
    /* hash_offset = exclusive scan of hash_count in u32 (mapper_mpe.c:250-255) */
    uint32_t *h_off = (uint32_t *)malloc((c->n_buckets + 1) * sizeof(uint32_t));
    if (!h_off) { free(rec); swm_ctx_destroy(c); return SWM_ENOMEM; }
    h_off[0] = 0;
    for (uint64_t i = 0; i < c->n_buckets; i++) h_off[i + 1] = h_off[i] + hash_count[i];
    /* buckets whose positions are not ascending (the serial scatter's
     * zeroed tail starts, Q23): their items are planned by k_plan_irreg */
    std::vector<uint32_t> irr;
    std::vector<uint64_t> descents, irr_bucket;
    for (uint64_t i = 0; i < c->n_buckets; i++) {
        const uint64_t a = h_off[i], b = (uint64_t)h_off[i] + hash_count[i];
        if (b > n_pos) break;
        bool any = false;
        for (uint64_t j = a + 1; j < b; j++)
            if (positions[j] < positions[j - 1]) { descents.push_back(j); any = true; }
        if (any) { irr.push_back((uint32_t)a); irr_bucket.push_back(i); }
    }
    /* with every offset below 2^31 (any index of a genome under 2^33 bases)
     * bit 31 of an irregular bucket's start entry flags it, so the seed
     * kernel learns it from the entry it gathers anyway */
    bool below31 = true;
    for (uint64_t i = 0; i <= c->n_buckets && below31; i++) below31 = h_off[i] < 0x80000000u;
    c->irr_flags = below31 && !irr.empty();
    /* the device keeps each k-mer's bucket next to its reverse complement's:
     * the two strands of a read select mostly the same long seeds, whose
     * 12-mers on the reverse strand are the forward ones' reverse
     * complements, so both fetch the same lines */
    uint32_t *d_order = (uint32_t *)malloc(c->n_buckets * sizeof(uint32_t));
    if (!d_order) { free(h_off); free(rec); swm_ctx_destroy(c); return SWM_ENOMEM; }
    uint64_t n_dev = pair_layout(hash_count, c->n_buckets, k, d_order, below31 ? 0x80000000ull : 0xffffff00ull);
    if (!n_dev) {
        /* the packed layout would not fit the offsets: pairs back to back */
        uint32_t acc = 0;
        for (uint64_t i = 0; i < c->n_buckets; i++) {
            const uint32_t rci = rc_bases((uint32_t)i, k);
            if (rci < i) continue;
            d_order[i] = acc;
            acc += hash_count[i];
            if (rci != i) { d_order[rci] = acc; acc += hash_count[rci]; }
        }
        n_dev = n_pos;
    }
    c->n_dev_pos = n_dev;
    for (uint32_t &x : irr) x = 0;
    for (size_t q = 0; q < irr_bucket.size(); q++) irr[q] = d_order[irr_bucket[q]];
    std::sort(irr.begin(), irr.end());
    if (c->irr_flags)
        for (uint64_t i : irr_bucket) d_order[i] |= 0x80000000u;
    int rc = SWM_OK;
    const size_t rec_bytes = (SWM_RECIPE_MAX_L + 1) * sizeof(swm_recipe_t);
    uint32_t *d_off = nullptr, *d_dev_off = nullptr;
    uint2 *d_pl_file = nullptr;
    const uint64_t tab_bytes = seed_table_entries(k) * sizeof(uint2);
    if (hipMalloc((void **)&d_off, (c->n_buckets + 1) * 4) != hipSuccess ||
        hipMalloc((void **)&d_dev_off, c->n_buckets * 4) != hipSuccess ||
        hipMalloc((void **)&d_pl_file, (n_pos + 1) * 8) != hipSuccess ||
        hipMalloc((void **)&c->d_seed_tab, tab_bytes) != hipSuccess ||
        hipMalloc((void **)&c->d_pl, (n_dev + 64) * 8) != hipSuccess ||
        hipMalloc((void **)&c->d_base_2bits, (n_2bits_words + 2 * SWM_WIN_PAIRS_PAD) * 4) != hipSuccess ||
        hipMalloc((void **)&c->d_recipe, rec_bytes) != hipSuccess ||
        (!irr.empty() && hipMalloc((void **)&c->d_irreg, irr.size() * 4) != hipSuccess))
        rc = SWM_ENOMEM;
    if (rc == SWM_OK) {
        if (hipMemcpy(d_off, h_off, (c->n_buckets + 1) * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(d_dev_off, d_order, c->n_buckets * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(c->d_seed_tab, 0, tab_bytes) != hipSuccess ||
            hipMemset(c->d_pl, 0, (n_dev + 64) * 8) != hipSuccess ||
            (n_pos && upload_pl(d_pl_file, positions, pos_lseeds, n_pos, descents) != SWM_OK) ||
            hipMemset(c->d_base_2bits, 0, (n_2bits_words + 2 * SWM_WIN_PAIRS_PAD) * 4) != hipSuccess ||
            hipMemcpy(c->d_base_2bits, base_2bits, n_2bits_words * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_recipe, rec, rec_bytes, hipMemcpyHostToDevice) != hipSuccess ||
            (!irr.empty() && hipMemcpy(c->d_irreg, irr.data(), irr.size() * 4, hipMemcpyHostToDevice) != hipSuccess))
            rc = SWM_EDEVICE;
        c->n_irreg = (uint32_t)irr.size();
    }
    if (rc == SWM_OK) {
        const unsigned nb = (unsigned)((c->n_buckets + 255) / 256);
        k_seed_table<<<nb, 256>>>(d_off, d_dev_off, c->n_buckets, k, c->d_seed_tab);
        if (n_pos) k_pl_permute<<<nb, 256>>>(d_off, d_dev_off, c->n_buckets, n_pos, n_dev, d_pl_file, c->d_pl);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = SWM_EDEVICE;
    }
    if (d_off) (void)hipFree(d_off);
    if (d_dev_off) (void)hipFree(d_dev_off);
    if (d_pl_file) (void)hipFree(d_pl_file);
    free(d_order);
    free(h_off);
    free(rec);
    if (rc != SWM_OK) { swm_ctx_destroy(c); return rc; }
    *out = c;
    return SWM_OK;
}

extern "C" uint64_t swm_seed_slot(uint32_t code, uint32_t k, uint32_t role) { return seed_slot(code, k, role & 1u); }

extern "C" void swm_ctx_destroy(swm_ctx_t *c)
{
    if (!c) return;
    swm_slot_destroy(c->slot0);
    if (c->d_seed_tab) (void)hipFree(c->d_seed_tab);
    if (c->d_pl) (void)hipFree(c->d_pl);
    if (c->d_base_2bits) (void)hipFree(c->d_base_2bits);
    if (c->d_recipe) (void)hipFree(c->d_recipe);
    if (c->d_irreg) (void)hipFree(c->d_irreg);
    delete c;
}

extern "C" int swm_last_stats(const swm_ctx_t *ctx, swm_stats_t *st)
{
    if (!ctx || !st) return SWM_EINVAL;
    *st = ctx->slot0->stats;
    return SWM_OK;
}

extern "C" int swm_ctx_set_option(swm_ctx_t *ctx, int opt, int64_t value)
{
    if (!ctx) return SWM_EINVAL;
    swm_opts &o = ctx->opt;
    switch (opt) {
    case SWM_OPT_WORK_QUEUE: if (value != 0 && value != 1) return SWM_EINVAL; o.work_queue = (uint32_t)value; break;
    case SWM_OPT_GRAPH: if (value != 0 && value != 1) return SWM_EINVAL; o.graph = (uint32_t)value; break;
    case SWM_OPT_HEAVY_CANDS: if (value < 1 || value > 65535) return SWM_EINVAL; o.heavy_cands = (uint32_t)value; break;
    case SWM_OPT_MYERS_BLOCK:
        if (value != 64 && value != 128 && value != 256) return SWM_EINVAL;
        o.myers_block = (uint32_t)value;
        break;
    case SWM_OPT_PLAN_HEAVY_WAVES: if (value < 1 || value > 32) return SWM_EINVAL; o.plan_heavy_waves = (uint32_t)value; break;
    case SWM_OPT_EMIT_HEAVY_WAVES: if (value < 1 || value > 64) return SWM_EINVAL; o.emit_heavy_waves = (uint32_t)value; break;
    case SWM_OPT_HISTOGRAM: if (value != 0 && value != 1) return SWM_EINVAL; o.histogram = (uint32_t)value; break;
    case SWM_OPT_MYERS_PACKED: if (value != 0 && value != 1) return SWM_EINVAL; o.myers_packed = (uint32_t)value; break;
    case SWM_OPT_GIANT_POS: if (value < 1 || value > (1 << 24)) return SWM_EINVAL; o.giant_pos = (uint32_t)value; break;
    case SWM_OPT_GIANT_CANDS: if (value < 1 || value > (1 << 24)) return SWM_EINVAL; o.giant_cands = (uint32_t)value; break;
    case SWM_OPT_SEED_GENERIC: if (value != 0 && value != 1) return SWM_EINVAL; o.seed_generic = (uint32_t)value; break;
    case SWM_OPT_PLAN_OVERLAP: if (value != 0 && value != 1) return SWM_EINVAL; o.plan_overlap = (uint32_t)value; break;
    case SWM_OPT_PLAN_LIGHT_WAVES: if (value < 8 || value > 64) return SWM_EINVAL; o.plan_light_waves = (uint32_t)value; break;
    default: return SWM_EINVAL;
    }
    return SWM_OK;
}

extern "C" int swm_ctx_get_option(const swm_ctx_t *ctx, int opt, int64_t *value)
{
    if (!ctx || !value) return SWM_EINVAL;
    const swm_opts &o = ctx->opt;
    switch (opt) {
    case SWM_OPT_WORK_QUEUE: *value = o.work_queue; break;
    case SWM_OPT_GRAPH: *value = o.graph; break;
    case SWM_OPT_HEAVY_CANDS: *value = o.heavy_cands; break;
    case SWM_OPT_MYERS_BLOCK: *value = o.myers_block; break;
    case SWM_OPT_PLAN_HEAVY_WAVES: *value = o.plan_heavy_waves; break;
    case SWM_OPT_EMIT_HEAVY_WAVES: *value = o.emit_heavy_waves; break;
    case SWM_OPT_HISTOGRAM: *value = o.histogram; break;
    case SWM_OPT_MYERS_PACKED: *value = o.myers_packed; break;
    case SWM_OPT_GIANT_POS: *value = o.giant_pos; break;
    case SWM_OPT_GIANT_CANDS: *value = o.giant_cands; break;
    case SWM_OPT_SEED_GENERIC: *value = o.seed_generic; break;
    case SWM_OPT_PLAN_OVERLAP: *value = o.plan_overlap; break;
    case SWM_OPT_PLAN_LIGHT_WAVES: *value = o.plan_light_waves; break;
    default: return SWM_EINVAL;
    }
    return SWM_OK;
}

static int check_map_args(const swm_ctx_t *ctx, uint64_t stride, uint32_t e)
{
    if (!ctx) return SWM_EINVAL;
    if (e > SWM_MAX_E) return SWM_EINVAL;              /* band < 32 bits (mapper_cpe.c:686) */
    if (stride < 3 || stride > 256) return SWM_EINVAL; /* -l rule, Q20 */
    return SWM_OK;
}

int swm_check_lens(const swm_ctx_t *ctx, uint64_t stride, const uint32_t *lens, uint64_t n)
{
    for (uint64_t i = 0; i < n; i++) {
        if ((uint64_t)lens[i] + 2 > stride) return SWM_EINVAL;                 /* kseq stride rule (Q20) */
        if (lens[i] / (ctx->k + SWM_SEED_STEP - 1) > 16) return SWM_EINVAL;   /* >16 long seeds */
    }
    return SWM_OK;
}

extern "C" int swm_map_batch_device(swm_ctx_t *ctx, const char *d_reads, uint64_t stride, const uint32_t *d_lens,
                                    uint64_t n, uint32_t e, swm_read_result_t *d_out)
{
    int rc = check_map_args(ctx, stride, e);
    if (rc) return rc;
    if (n && (!d_reads || !d_lens || !d_out)) return SWM_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    return swm_slot_run_device(ctx, ctx->slot0, d_reads, stride, d_lens, n, e, d_out, ctx->opt.graph != 0);
}

extern "C" int swm_map_batch_device_noalloc(swm_ctx_t *ctx, const char *d_reads, uint64_t stride,
                                            const uint32_t *d_lens, uint64_t n, uint32_t e,
                                            swm_read_result_t *d_out)
{
    int rc = check_map_args(ctx, stride, e);
    if (rc) return rc;
    if (n && (!d_reads || !d_lens || !d_out)) return SWM_EINVAL;
    if (n > SWM_NOALLOC_MAX_READS) return SWM_ENOSPC;
    if (!ctx->d_scratch) return SWM_EDEVICE;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    return run_noalloc(ctx, d_reads, stride, d_lens, n, e, d_out);
}

extern "C" int swm_map_batch(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens, uint64_t n,
                             uint32_t e, swm_read_result_t *out)
{
    int rc = check_map_args(ctx, stride, e);
    if (rc) return rc;
    if (n == 0) { memset(&ctx->slot0->stats, 0, sizeof(swm_stats_t)); return SWM_OK; }
    if (!reads || !lens || !out) return SWM_EINVAL;
    if ((rc = swm_check_lens(ctx, stride, lens, n))) return rc;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    return swm_slot_run_host(ctx, ctx->slot0, reads, stride, lens, n, n, e, out, ctx->opt.graph != 0);
}

extern "C" int swm_map_batch_records(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens,
                                     uint64_t n, uint32_t e, swm_read_result_t *out, swm_records_t *recs)
{
    if (!recs) return SWM_EINVAL;
    memset(recs, 0, sizeof(*recs));
    int rc = check_map_args(ctx, stride, e);
    if (rc) return rc;
    if (n == 0) { memset(&ctx->slot0->stats, 0, sizeof(swm_stats_t)); return SWM_OK; }
    if (!reads || !lens || !out) return SWM_EINVAL;
    if (2 * n >= 0xfffffff0ull) return SWM_EINVAL;
    if ((rc = swm_check_lens(ctx, stride, lens, n))) return rc;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    swm_slot *sl = ctx->slot0;
    sl->t_entry = host_ms();
    char *d_reads;
    uint32_t *d_lens;
    swm_read_result_t *d_out;
    if ((rc = sl->ws.get(swm_ws::IN_READS, &d_reads, n * stride + 4)) || (rc = sl->ws.get(swm_ws::IN_LENS, &d_lens, n)) ||
        (rc = sl->ws.get(swm_ws::OUT, &d_out, n)))
        return rc;
    hipStream_t st = sl->stream;
    SWM_HIP_CHECK(hipMemcpyAsync(d_reads, reads, n * stride, hipMemcpyHostToDevice, st));
    SWM_HIP_CHECK(hipMemcpyAsync(d_lens, lens, n * 4, hipMemcpyHostToDevice, st));
    if ((rc = run_synced(ctx, sl, d_reads, stride, d_lens, n, e, d_out, nullptr, 0, recs))) {
        swm_records_free(recs);
        return rc;
    }
    if (hipMemcpy(out, d_out, n * sizeof(swm_read_result_t), hipMemcpyDeviceToHost) != hipSuccess) {
        swm_records_free(recs);
        return SWM_EDEVICE;
    }
    return SWM_OK;
}

extern "C" void swm_records_free(swm_records_t *recs)
{
    if (!recs) return;
    free(recs->rec);
    free(recs->cigar);
    free(recs->block);
    memset(recs, 0, sizeof(*recs));
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
    hipStream_t st = ctx->slot0->stream;
    k_myers_explicit<<<blocks_for(n, 256), 256, 0, st>>>((uint32_t)n, d_read_words, d_lens, d_cand_read, d_win_start,
                                                         d_ref_len, ctx->d_base_2bits, e, d_err, d_loc);
    SWM_HIP_CHECK(hipGetLastError());
    SWM_HIP_CHECK(hipStreamSynchronize(st));
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

extern "C" int swm_host_alloc(void **p, uint64_t bytes)
{
    if (!p || !bytes) return SWM_EINVAL;
    *p = nullptr;
    return hipHostMalloc(p, bytes, hipHostMallocDefault) == hipSuccess ? SWM_OK : SWM_ENOMEM;
}

extern "C" int swm_host_free(void *p)
{
    if (!p) return SWM_OK;
    return hipHostFree(p) == hipSuccess ? SWM_OK : SWM_EDEVICE;
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

extern "C" int swm_copy_d2d(void *dst, const void *src, uint64_t bytes)
{
    if (bytes && (!dst || !src)) return SWM_EINVAL;
    return (!bytes || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess) ? SWM_OK : SWM_EDEVICE;
}

extern "C" int swm_dev_memset(void *dst, int value, uint64_t bytes)
{
    if (bytes && !dst) return SWM_EINVAL;
    return (!bytes || hipMemset(dst, value, bytes) == hipSuccess) ? SWM_OK : SWM_EDEVICE;
}

/* ------------------------------------------------------ VALU peak probe */

/* Integer VALU throughput: 8 independent chains per lane with per-lane
 * (VGPR) operands, each step an add followed by a 3-input boolean, so every
 * SIMD always has independent work.  Counts 3 VALU instructions per chain
 * step (add, xor, and-or). */
__global__ void __launch_bounds__(256) k_valu_probe(uint32_t iters, uint32_t seed, uint32_t *__restrict__ out)
{
    uint32_t x[8], y[8], z[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        x[j] = threadIdx.x * (j + 3) + seed;
        y[j] = (threadIdx.x ^ blockIdx.x) * 0x9e3779b9u + (uint32_t)j;
        z[j] = threadIdx.x * 0x85ebca6bu + (uint32_t)j * 7u;
    }
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = ((x[j] + y[j]) ^ z[j]) | (x[j] & y[(j + u) & 7]);
        }
    }
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) acc ^= x[j];
    if (acc == 0x12345678u) out[0] = acc;   /* keeps the chains live */
}

/* No counterpart in the reference: measures the chip's int32 VALU rate for
 * the Myers roofline (ops counted as lane operations per second). */
extern "C" int swm_valu_probe(double *ops_per_s)
{
    if (!ops_per_s) return SWM_EINVAL;
    uint32_t *d = nullptr;
    if (hipMalloc((void **)&d, 4) != hipSuccess) return SWM_ENOMEM;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { (void)hipFree(d); return SWM_EDEVICE; }
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const uint32_t blocks = (uint32_t)cus * 16, iters = 4096;
    k_valu_probe<<<blocks, 256>>>(16, 1u, d);                 /* warm-up */
    (void)hipEventRecord(a, 0);
    k_valu_probe<<<blocks, 256>>>(iters, 1u, d);
    (void)hipEventRecord(b, 0);
    int rc = hipEventSynchronize(b) == hipSuccess ? SWM_OK : SWM_EDEVICE;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    *ops_per_s = (double)blocks * 256.0 * iters * 8.0 * 8.0 * 3.0 / (ms * 1e-3);   /* add, xor, and_or */
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(d);
    return rc;
}

/* ---------------------------------------------------- gather rate probe */

/* Random 8-byte gathers, GP_PER per thread, all issued before any is used:
 * entry i of gather j is a hash of (thread, j) modulo n, so the lanes of a
 * wave (and the gathers of a thread) fall on distinct 128-byte lines and
 * every gather costs one line read. */
#define GP_PER 16
__device__ __forceinline__ uint32_t gp_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    return x ^ (x >> 16);
}
__global__ void __launch_bounds__(256) k_gather_probe(const uint2 *__restrict__ a, uint64_t n, uint32_t seed,
                                                      uint32_t *__restrict__ out)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    uint2 v[GP_PER];
#pragma unroll
    for (int j = 0; j < GP_PER; j++) {
        const uint64_t h = ((uint64_t)gp_hash(t * GP_PER + (uint32_t)j + seed) << 6) | (gp_hash(t ^ ((uint32_t)j * 977u)) & 63u);
        v[j] = a[h % n];
    }
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < GP_PER; j++) acc ^= v[j].x + v[j].y;
    if (acc == 0x9e3779b9u) out[t] = acc;   /* keeps the loads live */
}

/* No counterpart in the reference: the chip's rate of random 8-byte gathers
 * over an array of array_bytes (one line read each), in gathers per second
 * -- the ceiling of the gather-bound seed and bucket kernels, whose line
 * reads per launch rocprofv3 counts.  The best of three timed launches of
 * 64M gathers. */
extern "C" int swm_gather_probe(uint64_t array_bytes, double *gathers_per_s)
{
    if (!gathers_per_s || array_bytes < (1u << 20)) return SWM_EINVAL;
    const uint64_t n = array_bytes / 8;
    const uint32_t threads = 1u << 22;
    uint2 *a = nullptr;
    uint32_t *d = nullptr;
    if (hipMalloc((void **)&a, n * 8) != hipSuccess) return SWM_ENOMEM;
    if (hipMalloc((void **)&d, (uint64_t)threads * 4) != hipSuccess) { (void)hipFree(a); return SWM_ENOMEM; }
    int rc = hipMemset(a, 0x5a, n * 8) == hipSuccess ? SWM_OK : SWM_EDEVICE;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == SWM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = SWM_EDEVICE;
    float best = 0.0f;
    for (int rep = 0; rc == SWM_OK && rep < 4; rep++) {
        (void)hipEventRecord(e0, 0);
        k_gather_probe<<<threads / 256u, 256>>>(a, n, 0x1234u + (uint32_t)rep, d);
        (void)hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = SWM_EDEVICE; break; }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep && (best == 0.0f || ms < best)) best = ms;      /* the first launch warms up */
    }
    if (rc == SWM_OK) *gathers_per_s = (double)threads * GP_PER / (best * 1e-3);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(d);
    (void)hipFree(a);
    return rc;
}
