/*
 * swm_serve.hip -- serving loop (continuous batching over the resident
 * index) and the autotuner of the pipeline's tunables.
 *
 * The reference maps one file in double-buffered 10k-read batches: the host
 * fills buffer cur_block while the CPEs run the other (mapper_mpe.c:916-1106).
 * Here batches arrive from callers at any time.  Each of n_slots worker
 * threads owns a slot (a HIP stream with its own device workspace and
 * captured graphs) and takes the oldest waiting batch when the HBM budget
 * holds the batch's estimated workspace, so batches on different slots
 * overlap on the device while the index is shared read-only.
 */
#include "swm_internal.h"

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <string.h>
#include <thread>
#include <time.h>
#include <unordered_map>
#include <vector>

namespace {

double now_s()
{
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Batch sizes are padded to powers of two from 1024 reads, so a stream of
 * mixed sizes replays a handful of captured graphs per slot (the padding
 * reads have length 0 and cost almost nothing). */
uint64_t bucket(uint64_t n)
{
    uint64_t b = 1024;
    while (b < n) b <<= 1;
    return b;
}

struct Req {
    uint64_t ticket = 0;
    const char *reads = nullptr;
    const uint32_t *lens = nullptr;
    uint64_t n = 0, n_pad = 0, est = 0;
    swm_read_result_t *out = nullptr;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int status = SWM_OK;
    uint32_t slot = 0;
    bool done = false;
    uint64_t launch_reads = 0;
    uint32_t launch_batches = 0;
    swm_stats_t stats{};
};

}  // namespace

struct swm_server {
    swm_ctx_t *ctx = nullptr;
    swm_server_config_t cfg{};
    std::mutex mu;
    swm_gate gate;              /* the slots' runs take turns on the device; copies overlap */
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending;
    std::unordered_map<uint64_t, Req *> reqs;
    std::vector<std::thread> workers;
    std::vector<swm_slot *> slots;
    uint64_t next_ticket = 1, reserved = 0, peak = 0, budget = 0;
    uint64_t n_batches = 0, n_reads = 0, n_launches = 0;
    /* device bytes per padded read, learned from the slots' workspaces */
    double bytes_per_read = 0;
    uint64_t slot_max_pad[16] = {};
    bool stop = false;

    uint64_t estimate(uint64_t n_pad) const
    {
        return (uint64_t)((double)n_pad * bytes_per_read) + (1ull << 20);
    }

    void worker(uint32_t si)
    {
        (void)hipSetDevice(ctx->device);
        swm_slot *sl = slots[si];
        std::vector<Req *> take;
        std::vector<swm_segment_t> segs;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] {
                if (pending.empty()) return stop;
                return reserved + pending.front()->est <= budget;
            });
            if (pending.empty()) return;   /* stop and drained */
            /* the oldest batch (it fits, by the wait above), then every next
             * one that still fits the launch and the budget */
            take.clear();
            uint64_t tot = 0, est = 0;
            while (!pending.empty()) {
                Req *r = pending.front();
                if (!take.empty() && (cfg.max_launch == 0 || tot + r->n > cfg.max_launch)) break;
                if (!take.empty() && reserved + est + r->est > budget) break;
                pending.pop_front();
                take.push_back(r);
                tot += r->n;
                est += r->est;
            }
            reserved += est;
            peak = std::max(peak, reserved);
            const double ta = now_s();
            segs.clear();
            for (Req *r : take) {
                r->t_admit = ta;
                r->slot = si;
                segs.push_back(swm_segment_t{r->reads, r->lens, r->n, r->out});
            }
            lk.unlock();
            /* another waiting batch may fit on another slot */
            cv_work.notify_one();
            const uint64_t n_pad = cfg.graph ? bucket(tot) : tot;
            const int rc = swm_slot_run_segments(ctx, sl, segs.data(), (uint32_t)segs.size(), cfg.stride, n_pad,
                                                 cfg.e, cfg.graph != 0);
            const double t = now_s();
            const uint64_t ws = swm_slot_bytes(sl);
            lk.lock();
            for (Req *r : take) {
                r->status = rc;
                r->stats = *swm_slot_stats(sl);
                r->t_done = t;
                r->launch_reads = tot;
                r->launch_batches = (uint32_t)take.size();
                r->done = true;
                n_reads += r->n;
            }
            reserved -= est;
            slot_max_pad[si] = std::max(slot_max_pad[si], n_pad);
            bytes_per_read = std::max(bytes_per_read, (double)ws / (double)slot_max_pad[si]);
            n_batches += take.size();
            n_launches++;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" int swm_server_create(swm_ctx_t *ctx, const swm_server_config_t *cfg, swm_server_t **out)
{
    if (!ctx || !cfg || !out) return SWM_EINVAL;
    *out = nullptr;
    swm_server_config_t c = *cfg;
    if (c.n_slots == 0) c.n_slots = 4;
    if (c.max_batch == 0) c.max_batch = 65536;
    if (c.n_slots > 16 || c.e > SWM_MAX_E || c.stride < 3 || c.stride > 256 || c.max_batch > (1ull << 26) ||
        c.graph > 1 || (c.max_launch && (c.max_launch < c.max_batch || c.max_launch > (1ull << 26))))
        return SWM_EINVAL;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    swm_server *s = new (std::nothrow) swm_server();
    if (!s) return SWM_ENOMEM;
    s->ctx = ctx;
    s->cfg = c;
    if (c.hbm_budget) {
        s->budget = c.hbm_budget;
    } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete s; return SWM_EDEVICE; }
        s->budget = (uint64_t)(0.9 * (double)fr);
    }
    /* until a batch has run: the fixed per-read arrays (seeds, selections,
     * offsets, stride bytes in, result out) and a generous candidate share */
    s->bytes_per_read = 2048.0 + (double)c.stride + 64.0 * 32.0;
    for (uint32_t i = 0; i < c.n_slots; i++) {
        swm_slot *sl = swm_slot_create(ctx);
        if (!sl) { for (auto *x : s->slots) swm_slot_destroy(x); delete s; return SWM_EDEVICE; }
        if (c.graph) swm_slot_reserve(sl, bucket(std::max(c.max_batch, c.max_launch)));
        swm_slot_set_gate(sl, &s->gate);
        s->slots.push_back(sl);
    }
    try {
        for (uint32_t i = 0; i < c.n_slots; i++) s->workers.emplace_back(&swm_server::worker, s, i);
    } catch (...) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto &t : s->workers) t.join();
        for (auto *x : s->slots) swm_slot_destroy(x);
        delete s;
        return SWM_ENOMEM;
    }
    *out = s;
    return SWM_OK;
}

extern "C" int swm_server_submit(swm_server_t *s, const char *reads, const uint32_t *lens, uint64_t n,
                                 swm_read_result_t *out, uint64_t *ticket)
{
    if (!s || !ticket || n > s->cfg.max_batch || (n && (!reads || !lens || !out))) return SWM_EINVAL;
    int rc = swm_check_lens(s->ctx, s->cfg.stride, lens, n);
    if (rc) return rc;
    Req *r = new (std::nothrow) Req();
    if (!r) return SWM_ENOMEM;
    r->reads = reads;
    r->lens = lens;
    r->n = n;
    r->n_pad = s->cfg.graph ? bucket(n) : n;
    r->out = out;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->stop) { delete r; return SWM_EINVAL; }
        r->est = s->estimate(r->n_pad);
        if (r->est > s->budget) { delete r; return SWM_ENOMEM; }
        r->ticket = s->next_ticket++;
        s->reqs[r->ticket] = r;
        s->pending.push_back(r);
        *ticket = r->ticket;
    }
    s->cv_work.notify_one();
    return SWM_OK;
}

extern "C" int swm_server_wait(swm_server_t *s, uint64_t ticket, swm_batch_info_t *info)
{
    if (!s) return SWM_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return SWM_EINVAL;
    Req *r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    s->reqs.erase(ticket);   /* by key: submits during the wait may have rehashed the map */
    lk.unlock();
    if (info) {
        memset(info, 0, sizeof(*info));
        info->ticket = r->ticket;
        info->n_reads = r->n;
        info->t_submit = r->t_submit;
        info->t_admit = r->t_admit;
        info->t_done = r->t_done;
        info->status = r->status;
        info->slot = r->slot;
        info->est_bytes = r->est;
        info->launch_reads = r->launch_reads;
        info->launch_batches = r->launch_batches;
        info->stats = r->stats;
    }
    const int rc = r->status;
    delete r;
    return rc;
}

/* Maps the sample (repeated to size) once per slot at every batch bucket up
 * to max_batch, before any traffic: each slot learns its capacities and
 * captures its graphs, so served batches never pay for a capture.  Must not
 * overlap submissions. */
extern "C" int swm_server_warmup(swm_server_t *s, const char *reads, const uint32_t *lens, uint64_t n)
{
    if (!s || !reads || !lens || n == 0) return SWM_EINVAL;
    int rc = swm_check_lens(s->ctx, s->cfg.stride, lens, n);
    if (rc) return rc;
    if (hipSetDevice(s->ctx->device) != hipSuccess) return SWM_EDEVICE;
    const uint64_t most = std::max(s->cfg.max_batch, s->cfg.max_launch);
    const uint64_t top = s->cfg.graph ? bucket(most) : most;
    std::vector<char> rb;
    std::vector<uint32_t> lb;
    std::vector<swm_read_result_t> ob;
    try {
        rb.resize(top * s->cfg.stride);
        lb.resize(top);
        ob.resize(top);
    } catch (...) {
        return SWM_ENOMEM;
    }
    for (uint64_t i = 0; i < top; i++) {
        memcpy(&rb[i * s->cfg.stride], reads + (i % n) * s->cfg.stride, s->cfg.stride);
        lb[i] = lens[i % n];
    }
    std::vector<uint64_t> sizes;
    if (s->cfg.graph) {
        for (uint64_t b = 1024; b <= top; b <<= 1) sizes.push_back(b);
        sizes.insert(sizes.begin(), top);   /* the largest first: one growth, then none */
    } else {
        sizes.push_back(top);
    }
    for (swm_slot *sl : s->slots) {
        for (uint64_t b : sizes) {
            /* twice: the first run of a slot sizes its capacities (synced), the next captures */
            for (int k = 0; k < 2 && rc == SWM_OK; k++)
                rc = swm_slot_run_host(s->ctx, sl, rb.data(), s->cfg.stride, lb.data(), b, b, s->cfg.e, ob.data(),
                                       s->cfg.graph != 0);
        }
    }
    if (rc == SWM_OK) {
        std::lock_guard<std::mutex> g(s->mu);
        for (size_t i = 0; i < s->slots.size(); i++) {
            s->slot_max_pad[i] = std::max(s->slot_max_pad[i], top);
            s->bytes_per_read = std::max(s->bytes_per_read, (double)swm_slot_bytes(s->slots[i]) / (double)top);
        }
    }
    return rc;
}

extern "C" int swm_server_totals(const swm_server_t *s, uint64_t *batches, uint64_t *reads, uint64_t *peak_bytes,
                                 uint64_t *budget)
{
    if (!s) return SWM_EINVAL;
    std::lock_guard<std::mutex> g(const_cast<swm_server *>(s)->mu);
    if (batches) *batches = s->n_batches;
    if (reads) *reads = s->n_reads;
    if (peak_bytes) *peak_bytes = s->peak;
    if (budget) *budget = s->budget;
    return SWM_OK;
}

extern "C" void swm_server_destroy(swm_server_t *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    for (auto *x : s->slots) swm_slot_destroy(x);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
}

/* ------------------------------------------------------------- autotuner */

namespace {

int set_opt(swm_ctx_t *ctx, int opt, int64_t v) { return swm_ctx_set_option(ctx, opt, v); }

}  // namespace

/* Coordinate search: each tunable in turn over its candidate values, the
 * others held at the best found so far; a setting is timed as the median of
 * reps wall-clock runs of the whole pipeline on device-resident input. */
extern "C" int swm_autotune(swm_ctx_t *ctx, const char *reads, uint64_t stride, const uint32_t *lens, uint64_t n,
                            uint32_t e, uint32_t reps, int64_t *chosen)
{
    if (!ctx || !reads || !lens || n == 0 || reps == 0 || reps > 64 || e > SWM_MAX_E || stride < 3 || stride > 256)
        return SWM_EINVAL;
    int rc = swm_check_lens(ctx, stride, lens, n);
    if (rc) return rc;
    if (hipSetDevice(ctx->device) != hipSuccess) return SWM_EDEVICE;
    char *d_reads = nullptr;
    uint32_t *d_lens = nullptr;
    swm_read_result_t *d_out = nullptr;
    if (hipMalloc((void **)&d_reads, n * stride) != hipSuccess || hipMalloc((void **)&d_lens, n * 4) != hipSuccess ||
        hipMalloc((void **)&d_out, n * sizeof(swm_read_result_t)) != hipSuccess) {
        if (d_reads) (void)hipFree(d_reads);
        if (d_lens) (void)hipFree(d_lens);
        return SWM_ENOMEM;
    }
    rc = SWM_OK;
    if (hipMemcpy(d_reads, reads, n * stride, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_lens, lens, n * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = SWM_EDEVICE;

    auto time_it = [&](double *ms) -> int {
        /* one untimed run (captures the graph, grows the workspace) */
        int r = swm_slot_run_device(ctx, ctx->slot0, d_reads, stride, d_lens, n, e, d_out, ctx->opt.graph != 0);
        if (r) return r;
        std::vector<double> t;
        for (uint32_t i = 0; i < reps; i++) {
            const double t0 = now_s();
            r = swm_slot_run_device(ctx, ctx->slot0, d_reads, stride, d_lens, n, e, d_out, ctx->opt.graph != 0);
            if (r) return r;
            t.push_back(1e3 * (now_s() - t0));
        }
        std::sort(t.begin(), t.end());
        *ms = t[t.size() / 2];
        return SWM_OK;
    };

    struct Knob { int opt; std::vector<int64_t> vals; };
    const Knob knobs[] = {
        {SWM_OPT_WORK_QUEUE, {1, 0}},
        {SWM_OPT_MYERS_PACKED, {1, 0}},
        {SWM_OPT_MYERS_BLOCK, {256, 128, 64}},
        {SWM_OPT_HEAVY_CANDS, {64, 32, 128, 256, 512}},
        {SWM_OPT_PLAN_HEAVY_WAVES, {32, 8, 16}},
        {SWM_OPT_PLAN_OVERLAP, {0, 1}},
        {SWM_OPT_PLAN_LIGHT_WAVES, {32, 64}},
        {SWM_OPT_EMIT_HEAVY_WAVES, {32, 8, 16, 64}},
    };
    for (const Knob &kb : knobs) {
        if (rc) break;
        int64_t best_v = 0;
        (void)swm_ctx_get_option(ctx, kb.opt, &best_v);
        double best_ms = 0;
        bool have = false;
        for (int64_t v : kb.vals) {
            if ((rc = set_opt(ctx, kb.opt, v))) break;
            double ms;
            if ((rc = time_it(&ms))) break;
            if (!have || ms < best_ms) { best_ms = ms; best_v = v; have = true; }
        }
        if (!rc) rc = set_opt(ctx, kb.opt, best_v);
    }
    if (!rc && chosen)
        for (int o = 1; o <= SWM_N_OPTIONS; o++) (void)swm_ctx_get_option(ctx, o, &chosen[o - 1]);
    (void)hipFree(d_reads);
    (void)hipFree(d_lens);
    (void)hipFree(d_out);
    return rc;
}
