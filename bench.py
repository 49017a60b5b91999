#!/usr/bin/env python3
"""Benchmark of the mapping hot path on one MI355X.

Workload (default): 256 Mbp uniform-random genome with 5% inserted repeat
families (300 bp units x 500 copies at 10% divergence), 12-mer index; 1M
reads of 150 bp sampled from both strands with 2% substitutions and 0.5%
1-3 bp indels; e = 5.  A step maps the whole 1M-read batch (reads resident
in HBM) through seed lookup, candidate batches, banded Myers and record
emission.  Prints one JSON line.
"""
import argparse
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MYERS_OPS_PER_STEP = 20          # int32 ops per band step (X 1, D0 4, HN 1, HP 3, >>1 1, VN 1, VP 3, tally 2, Peq 4)
VALU_PEAK_OPS = 256 * 4 * 16 * 2.4e9   # CUs x SIMDs x lanes x clock = 39.3e12 int32 ops/s
HBM_PEAK_GBS = 8000.0

PMC_FILES = {"repeat5": "map_repeat5_1Mx150_e5_pmc_bytes.txt"}
REQ_FILES = {"repeat5": "map_repeat5_1Mx150_e5_pmc_requests.txt"}
ISSUE_FILE = "map_valu_issue_skew_repeat5.txt"


def pmc_traffic(config, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (read_MB + write_MB columns, already
    scaled by the copy-kernel calibration), or None."""
    name = PMC_FILES.get(config)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if f and f[0] == kernel and len(f) >= 5:
            return {"bytes": int((float(f[3]) + float(f[4])) * 1e6), "source": "profiles/" + name}
    return None


def pmc_line_reads(config, kernel):
    """Memory read requests (128-byte lines) per launch of a kernel from the
    committed rocprofv3 TCC_EA0_RDREQ summary of this workload, or None."""
    name = REQ_FILES.get(config)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if f and f[0] == kernel and len(f) >= 3:
            try:
                return int(f[1])
            except ValueError:
                return None
    return None


def gather_ceiling(line_reads, ms, probe):
    """A gather-bound kernel against the measured random-gather rate: its
    line reads per second over the probe's (one line per 8-byte gather)."""
    if not line_reads or not ms or not probe:
        return None
    rate = line_reads / (ms / 1e3)
    return {"line_reads": line_reads, "line_reads_per_s": round(rate / 1e9, 2) * 1e9,
            "probe_gathers_per_s": round(probe / 1e9, 2) * 1e9, "frac": round(rate / probe, 4),
            "unit": "line reads/s", "source": "profiles/" + REQ_FILES.get("repeat5", "")}


def valu_issue(config, kernel):
    """A kernel's VALU issue fraction on this workload from the committed
    rocprofv3 summary (SQ_INSTS_VALU x 4 cycles over SIMDs x the dispatch's
    GRBM_GUI_ACTIVE cycles, all from one run), or None."""
    path = os.path.join(ROOT, "profiles", ISSUE_FILE)
    if not os.path.exists(path):
        return None
    sect = None
    for line in open(path):
        if line.startswith("## "):
            sect = line[3:].split()[0]
        f = line.split()
        if sect == config and f and f[0] == kernel and len(f) >= 4:
            try:
                return {"valu_issue": float(f[2]), "source": "profiles/" + ISSUE_FILE}
            except ValueError:
                return None
    return None


def host_cpu():
    """The host's CPU model and core counts (lscpu, run as a child and waited
    for) and the threads this process may use."""
    info = {"affinity_threads": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "CPU(s)", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def child_processes():
    """Processes whose parent is this one (none should be left at exit)."""
    me, kids = os.getpid(), []
    for d in os.listdir("/proc"):
        if d.isdigit():
            try:
                with open(f"/proc/{d}/stat") as f:
                    if int(f.read().rsplit(")", 1)[1].split()[1]) == me:
                        kids.append(int(d))
            except (OSError, ValueError, IndexError):
                pass
    return kids


def traffic_rate(traffic, ms):
    """The memory system's side of a gather-bound kernel: the HBM bytes it
    moves per launch (whole lines) over its launch time, against the peak."""
    if not traffic or not ms:
        return None
    gbs = traffic["bytes"] / (ms / 1e3) / 1e9
    return {"gbs": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}

FORMS = {
    # one launch per stage, one thread per (read, strand) for every item
    "plain": {1: 0, 2: 0},
    # heavy items through device work queues (persistent single-wave plan
    # kernel, cooperative compaction and emission); stage-timed, one stream
    # (the heavy-item plan kernels after the light-item one), so that every
    # stage and the roofline kernels are timed running alone
    "queue": {1: 1, 2: 0, 12: 0},
    # the work-queue pipeline with device-side counts and fixed capacities,
    # captured once and replayed as a hipGraph (no host round trip per batch)
    "graph": {1: 1, 2: 1, 12: 0},
}


def rooflines(swm, st_all, lens, config, form):
    """Rooflines from a stage-timed form (HIP events around each stage, and
    around k_plan_pair, on the context stream): the seed lookup (k_seed,
    HBM-bound random gathers; algorithmic bytes = read bytes + 8 B per
    12-mer queried, 4 per long seed, both strands), the light-item plan
    kernel (k_plan_pair, HBM-bound random bucket fetches; 6 B per bucket
    entry it fetched), the whole plan stage, and banded Myers (VALU-bound,
    20 int32 ops per band step).  `roofline` is the one of k_seed and
    k_plan_pair with the longer launch: the dominant kernel."""
    k = 12
    lv = lens.astype(np.int64)
    ok = (lv + 2 <= 256) & (lv // (k + 3) <= 16)
    seed_bytes = int(lv[ok].sum() + 2 * 4 * 8 * (lv[ok] // (k + 3)).sum())
    roof_seed = roof_wave = roof_plan = roof_myers = None
    # the chip's random 8-byte gather rate over an array the size of the
    # bucket array: the ceiling of the two gather-bound kernels
    probe = swm.gather_peak(512 << 20) if config in REQ_FILES else None
    ms_seed = statistics.mean(s["ms_seed"] for s in st_all)
    if ms_seed > 0:
        traffic = pmc_traffic(config, "k_seed_fixed<12u>") or pmc_traffic(config, "k_seed")
        ach = seed_bytes / (ms_seed / 1e3) / 1e9
        roof_seed = {"bound": "hbm", "kernel": "k_seed_fixed<12> (k_seed for other k)", "form": form,
                     "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_rate": traffic_rate(traffic, ms_seed),
                     "algorithmic_bytes_per_launch": seed_bytes, "avg_launch_ms": round(ms_seed, 3),
                     "gather": gather_ceiling(pmc_line_reads(config, "k_seed_fixed<12u>"), ms_seed, probe),
                     "note": "random 8-byte gathers into the seed table: a 128-byte line serves two "
                             "consecutive 12-mers of a read and their reverse complements, so HBM traffic "
                             "(rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated on a copy) is still several "
                             "times the algorithmic bytes"}
    ms_wave = statistics.mean(s.get("ms_plan_light", 0.0) for s in st_all)
    lpos = st_all[-1].get("n_light_positions", 0)
    if ms_wave > 0 and lpos:
        traffic = pmc_traffic(config, "k_plan_pair<32u>") or pmc_traffic(config, "k_plan_pair")
        ach = 6 * lpos / (ms_wave / 1e3) / 1e9
        roof_wave = {"bound": "hbm", "kernel": "k_plan_pair (both strands of a read per wave)", "form": form,
                     "achieved": round(ach, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_rate": traffic_rate(traffic, ms_wave),
                     "algorithmic_bytes_per_launch": 6 * lpos, "avg_launch_ms": round(ms_wave, 3),
                     "gather": gather_ceiling(pmc_line_reads(config, "k_plan_pair<32u>") or pmc_line_reads(config, "k_plan_pair"), ms_wave, probe),
                     "note": "random bucket fetches (about 4 entries of 8 bytes each) into the 512 MB "
                             "bucket array, both strands of a read in one load: each still costs a whole line"}
    ms_plan = statistics.mean(s["ms_plan"] for s in st_all)
    npos = st_all[-1]["n_positions"]
    if ms_plan > 0 and npos:
        ach = 6 * npos / (ms_plan / 1e3) / 1e9
        roof_plan = {"bound": "hbm", "kernel": "plan stage (k_plan_pair, heavy plan kernels, large-batch sort)",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": 6 * npos,
                     "avg_launch_ms": round(ms_plan, 3)}
    sum_l_ver = st_all[-1]["n_band_steps"]
    ms_myers = statistics.mean(s["ms_myers"] for s in st_all)
    if sum_l_ver and ms_myers > 0:
        # peak: the larger of the 64-lanes-per-CU-per-clock figure and the
        # on-box add/xor throughput probe (the probe is what the chip does)
        measured = swm.valu_peak()
        peak = max(VALU_PEAK_OPS, measured)
        ach = MYERS_OPS_PER_STEP * sum_l_ver / (ms_myers / 1e3)
        roof_myers = {"bound": "valu", "kernel": "k_myers2 (e <= 7; k_myers above)", "form": form,
                      "achieved": round(ach / 1e12, 3), "peak": round(peak / 1e12, 2), "unit": "Tops/s (int32)",
                      "frac": round(ach / peak, 4), "traffic": None,
                      "peak_probe_tops": round(measured / 1e12, 2),
                      "peak_64lane_tops": round(VALU_PEAK_OPS / 1e12, 2),
                      "issue": valu_issue(config, "k_myers2"),
                      "note": "the kernel also runs the Hamming check of each passing candidate (its window "
                              "reloaded from the cache), which the count leaves out; "
                              "achieved counts 20 algorithmic int32 ops per band step of each candidate; with "
                              "equal read lengths two candidates share a lane (16-bit halves) and gfx950's "
                              "3-input boolean ops fuse pairs, so frac is not capped at 1; 'issue' is the "
                              "kernel's measured VALU instruction issue fraction (rocprofv3 counters of one "
                              "run), which is capped at 1",
                      "algorithmic_ops_per_launch": MYERS_OPS_PER_STEP * sum_l_ver,
                      "avg_launch_ms": round(ms_myers, 3)}
    cands = [x for x in (roof_seed, roof_wave) if x]
    roof = max(cands, key=lambda x: x["avg_launch_ms"]) if cands else None
    return {"roofline": roof, "roofline_seed": roof_seed, "roofline_plan_light": roof_wave,
            "roofline_plan": roof_plan, "roofline_myers": roof_myers}


FIELDS = ("count", "pos", "strand", "dist")


def sample_check(orc, ix, reads, lens, e, got, seconds, heavy=0):
    """The oracle on a sample of a batch -- its first reads, as many as take
    about `seconds` on every host core, plus the `heavy` reads with the most
    records -- and whether the GPU's results are identical there."""
    ref = orc.OracleRef(12, ix.hash_count, ix.positions, ix.pos_lseeds, ix.base_2bits)
    threads = len(os.sched_getaffinity(0))
    n = lens.size
    probe = min(n, 2000)
    t = time.perf_counter()
    ref.map(reads[:probe], lens[:probe], e, threads)
    dt = time.perf_counter() - t
    ns = int(min(n, max(probe, probe * seconds / max(dt, 1e-3))))
    idx = np.arange(ns)
    if heavy:
        idx = np.union1d(idx, np.argsort(got["count"], kind="stable")[::-1][:heavy])
    t = time.perf_counter()
    out, _ = ref.map(reads[idx], lens[idx], e, threads)
    dt = time.perf_counter() - t
    same = all(np.array_equal(out[f], got[f][idx]) for f in FIELDS)
    what = f"first {ns}" + (f" + the {heavy} with the most records" if heavy else "")
    return {"gpu_results_identical": bool(same),
            "sample": f"{idx.size} reads ({what}) of {n}, oracle on {threads} threads, {dt:.1f}s"}


def measure_config(swm, orc, name, cfg, g, ix, forms, args, log, e, tune=False, heavy=0, check_s=3.0):
    """One configuration on the resident index: its reads, the given forms
    timed (after the autotuner on a 100k sample if tune), rooflines from one
    stage-timed (work-queue) step, and an oracle check on a sample."""
    from swmapper_amd.configs import READS_SEED
    reads, lens, _, _ = swm.synth_reads(g, cfg["reads"], len_min=cfg["lmin"], len_max=cfg["lmax"],
                                        err_min=cfg["err"][0], err_max=cfg["err"][1], indel_share=0.2,
                                        seed=READS_SEED, stride=256)
    n = lens.size
    m = swm.Mapper.from_index(ix)
    d_reads = swm.DeviceBuffer.from_array(reads)
    d_lens = swm.DeviceBuffer.from_array(lens)
    d_out = swm.DeviceBuffer(n * swm.RESULT_DTYPE.itemsize).zero()

    def step():
        m.map_device(d_reads.ptr, 256, d_lens.ptr, n, e, d_out.ptr)
        return m.stats()

    tuned = m.autotune(reads[:min(n, 100_000)], lens[:min(n, 100_000)], e, reps=3) if tune else None
    out = {"workload": cfg["desc"].rsplit("; e=", 1)[0] + f"; e={e}", "reads": n, "e": e, "tuned": tuned,
           "forms": {}}
    first = None
    for f in forms:
        elapsed, lat, st_all = time_form(m, FORMS[f], step, args.config_steps, 1)
        res = d_out.to_array(swm.RESULT_DTYPE)
        if first is None:
            first = res
        nv = st_all[-1]["n_verifications"]
        out["forms"][f] = {"ms_per_step": round(elapsed * 1e3 / args.config_steps, 3),
                           "verifications_per_s": round(nv * args.config_steps / elapsed, 1),
                           "reads_per_s": round(n * args.config_steps / elapsed, 1),
                           "identical_results": bool(np.array_equal(res, first))}
    best = min(out["forms"], key=lambda f: out["forms"][f]["ms_per_step"])
    out.update({"form": best, **{k: out["forms"][best][k] for k in ("ms_per_step", "verifications_per_s",
                                                                     "reads_per_s")}})
    _, _, st_q = time_form(m, FORMS["queue"], step, 1, 0)
    roofs = rooflines(swm, st_q, lens, name, "queue")
    stages = {k: round(st_q[-1][k], 3) for k in ("ms_seed", "ms_plan", "ms_compact", "ms_myers", "ms_emit")}
    out.update({**roofs, "stage_ms": stages,
                "dominant_stage": max(stages, key=stages.get), "verifications_per_step": st_q[-1]["n_verifications"],
                "giant_items": st_q[-1].get("n_giant_items"), "giant_reads": st_q[-1].get("n_giant_reads")})
    out.update(sample_check(orc, ix, reads, lens, e, first, check_s, heavy))
    log(f"config {name} e={e}: {best} {out['ms_per_step']} ms/step, identical on sample: "
        f"{out['gpu_results_identical']}")
    m.close()
    for b in (d_reads, d_lens, d_out):
        b.free()
    return out


def time_form(m, opts, step, steps, warmup):
    for k, v in opts.items():
        m.set_option(k, v)
    # the graph form's first call runs synced to size its capacities and the
    # second captures the graph: both before the timed steps
    if opts.get(2):
        warmup = max(warmup, 2)
    for _ in range(warmup):
        step()
    lat, st_all = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        st_all.append(step())
        lat.append((time.perf_counter() - ts) * 1e3)
    return time.perf_counter() - t0, lat, st_all


def serve_bench(swm, m, g, e, seconds, log, slots=2, graph=True, merge=262144):
    """Continuous batching on the resident index: batches of U[1k, 64k] reads
    with lengths U[50, 250] from page-locked memory, Poisson arrivals at
    0.3 / 0.6 / 0.9 of the measured saturation rate."""
    import threading
    pool_n = 1 << 18
    pr, pl, _, _ = swm.synth_reads(g, pool_n, len_min=50, len_max=250, err_min=0.01, err_max=0.04,
                                   indel_share=0.2, seed=4242, stride=256)
    pin_r = swm.PinnedArray(pr.shape, np.uint8)
    pin_l = swm.PinnedArray(pl.shape, np.uint32)
    pin_r.a[:] = pr
    pin_l.a[:] = pl
    offline = m.map(pr, pl, e)          # the same reads mapped offline, for the bit-exact check
    rng = np.random.default_rng(99)

    def batches(k):
        out = []
        for _ in range(k):
            n = int(rng.integers(1024, 65537))
            s = int(rng.integers(0, pool_n - n))
            out.append((s, n))
        return out

    srv = swm.Server(m, e=e, stride=256, n_slots=slots, graph=graph, max_launch=merge)
    exact = True

    def check(s, n, got):
        nonlocal exact
        ref = offline[s:s + n]
        for f in ("count", "pos", "strand", "dist"):
            if not np.array_equal(got[f], ref[f]):
                exact = False

    # capacities and graphs of every slot up front, then a few batches, then saturate
    srv.warmup(pr[:8192], pl[:8192])
    for s, n in batches(8):
        t = srv.submit(pin_r.a[s:s + n], pin_l.a[s:s + n])
        got, _ = srv.wait(t)
        check(s, n, got)
    sat = batches(48)
    t0 = time.monotonic()
    ts = [srv.submit(pin_r.a[s:s + n], pin_l.a[s:s + n]) for s, n in sat]
    done, service, gpu, henq, hwait, lb = [], [], [], [], [], []
    for t, (s, n) in zip(ts, sat):
        got, info = srv.wait(t)
        check(s, n, got)
        done.append(info["t_done"])
        service.append((info["t_done"] - info["t_admit"]) * 1e3)
        gpu.append(info["stats"]["ms_total"])
        henq.append(info["stats"]["ms_host_enqueue"])
        hwait.append(info["stats"]["ms_host_wait"])
        lb.append(info["launch_batches"])
    max_rps = sum(n for _, n in sat) / (max(done) - t0)
    sat_info = {"reads_per_s": round(max_rps, 1), "batches": len(sat),
                "service_ms_mean": round(float(np.mean(service)), 3),
                "device_ms_mean": round(float(np.mean(gpu)), 3),
                "host_enqueue_ms_mean": round(float(np.mean(henq)), 3),
                "host_wait_ms_mean": round(float(np.mean(hwait)), 3),
                "batches_per_launch_mean": round(float(np.mean(lb)), 2)}
    log(f"serving saturation: {sat_info}")
    mean_n = (1024 + 65536) / 2
    loads = []
    for load in (0.3, 0.6, 0.9):
        rate = load * max_rps / mean_n                      # batches per second
        plan = []
        t = 0.0
        while t < seconds:
            t += rng.exponential(1.0 / rate)
            plan.append(t)
        work = batches(len(plan))
        results = []

        def waiter(tickets):
            for (tk, arr, s, n) in tickets:
                got, info = srv.wait(tk)
                results.append((info["t_done"] - arr, n, info["t_done"], info["t_done"] - info["t_admit"],
                                info["stats"]["ms_total"], info["launch_batches"]))
                check(s, n, got)

        tickets = []
        base = time.monotonic() + 0.05
        for (s, n), at in zip(work, plan):
            arr = base + at
            dt = arr - time.monotonic()
            if dt > 0:
                time.sleep(dt)
            tickets.append((srv.submit(pin_r.a[s:s + n], pin_l.a[s:s + n]), arr, s, n))
        th = threading.Thread(target=waiter, args=(tickets,))
        th.start()
        th.join()
        lat = np.array([r[0] for r in results]) * 1e3
        nreads = sum(r[1] for r in results)
        span = max(r[2] for r in results) - base
        loads.append({"offered": load, "offered_reads_per_s": round(load * max_rps, 1),
                      "reads_per_s": round(nreads / span, 1), "batches": len(results),
                      "p50_ms": round(float(np.percentile(lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat, 99)), 3),
                      "service_ms_mean": round(float(np.mean([r[3] for r in results])) * 1e3, 3),
                      "device_ms_mean": round(float(np.mean([r[4] for r in results])), 3),
                      "batches_per_launch_mean": round(float(np.mean([r[5] for r in results])), 2)})
        log(f"serving load {load}: {loads[-1]}")
    tot = srv.totals()
    srv.close()
    pin_r.free()
    pin_l.free()
    return {"workload": "batches of U[1k,64k] reads, lengths U[50,250], error U[1%,4%], from page-locked "
                        f"memory; Poisson arrivals; {slots} slots (streams), {'graph' if graph else 'synced'} form, "
                        f"FIFO admission by HBM budget, waiting batches merged up to {merge} reads per launch",
            "max_reads_per_s": round(max_rps, 1), "saturation": sat_info, "loads": loads,
            "bit_exact_vs_offline": exact,
            "peak_admitted_bytes": tot["peak_bytes"], "hbm_budget_bytes": tot["budget"]}


def records_bench(swm, m, reads, lens, e, ref_out, n_bases, log):
    """The records path (every record with its CIGAR text, host buffers in
    and out) on the whole batch, then the SAM writer (gzip level 1, as the
    reference's <out>.sam.gz) on the first 100k reads' records."""
    import tempfile
    n = lens.size
    m.map_records(reads[:1000], lens[:1000], e)                 # grows the record buffers
    t0 = time.perf_counter()
    res, recs, text = m.map_records(reads, lens, e)
    dt = time.perf_counter() - t0
    same = all(np.array_equal(res[f], ref_out[f]) for f in ("count", "pos", "strand", "dist"))
    k = min(n, 100_000)
    sub = recs[recs["read"] < k]
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        with open(os.path.join(td, "ix.ref_info"), "wb") as f:
            f.write(b"%d\n9\nsynthetic\n" % n_bases)
        names = [b"r%d" % i for i in range(k)]
        quals = np.full((k, reads.shape[1]), ord("I"), np.uint8)
        path = os.path.join(td, "out.sam.gz")
        t1 = time.perf_counter()
        w = swm.SamWriter(path, os.path.join(td, "ix"), gzip_level=1)
        w.write(names, reads[:k], quals, lens[:k], sub, text)
        w.close()
        ds = time.perf_counter() - t1
        gz_bytes = os.path.getsize(path)
        path = os.path.join(td, "out.sam")
        t2 = time.perf_counter()
        w = swm.SamWriter(path, os.path.join(td, "ix"), gzip_level=0)
        w.write(names, reads[:k], quals, lens[:k], sub, text)
        w.close()
        dp = time.perf_counter() - t2
        sam_bytes = os.path.getsize(path)
    out = {"entry": "swm_map_batch_records", "ms": round(dt * 1e3, 3), "reads_per_s": round(n / dt, 1),
           "records": int(recs.size), "cigar_bytes": len(text), "identical_results": same,
           "note": "wall time: host reads in (pageable), synced pipeline, record compaction, records + CIGAR "
                   "text back to host memory",
           "sam": {"reads": k, "records": int(sub.size), "sam_bytes": sam_bytes, "gzip_bytes": gz_bytes,
                   "gzip1_ms": round(ds * 1e3, 1), "plain_ms": round(dp * 1e3, 1),
                   "plain_mb_per_s": round(sam_bytes / dp / 1e6, 1), "gzip1_mb_per_s": round(sam_bytes / ds / 1e6, 1)}}
    log(f"records: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="repeat5", choices=("repeat5", "small", "uneven", "skew"))
    ap.add_argument("--e", type=int, default=None, help="override the configuration's error bound")
    ap.add_argument("--forms", default="plain,queue,graph", help="pipeline forms to time (plain, queue, graph)")
    ap.add_argument("--autotune", action="store_true", help="autotune the tunables on a 100k-read sample first")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="length of each serving load (0: no serving)")
    ap.add_argument("--serve-slots", type=int, default=2, help="serving slots (streams)")
    ap.add_argument("--serve-merge", type=int, default=262144,
                    help="reads per serving launch when waiting batches are merged (0: one batch per launch)")
    ap.add_argument("--serve-form", default="graph", choices=("graph", "synced"))
    ap.add_argument("--serve-wq", type=int, default=None, choices=(0, 1),
                    help="work queues in the serving launches (default: those of the fastest offline form)")
    ap.add_argument("--hw-queues", type=int, default=8,
                    help="GPU_MAX_HW_QUEUES for this process when not set by the caller above 4: the serving "
                         "slots' streams each get a hardware queue")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-records", action="store_true", help="skip the records + SAM measurement")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="set a tunable (swm_ctx_set_option by name, e.g. seed_cache=0) before timing")
    ap.add_argument("--index-only", type=int, default=0, metavar="K",
                    help="only build the configuration's index K times and print the build times (for profiling)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="one 256 MiB device-to-device copy first (calibrates HBM byte counters)")
    ap.add_argument("--configs", default="small,uneven,skew",
                    help="the other configurations measured after the headline (comma list; empty: none)")
    ap.add_argument("--config-steps", type=int, default=2, help="timed steps per form in each other configuration")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(mapping shards by reads across GPUs; see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    # before the HIP runtime starts (it reads the variable once)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) <= 4 and 1 <= args.hw_queues <= 32:
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    import swmapper_amd as swm
    from swmapper_amd.configs import CONFIGS, GENOME_SEED, READS_SEED
    cfg = dict(CONFIGS[args.config])
    if args.e is not None:
        cfg["e"] = args.e
        cfg["desc"] = cfg["desc"].rsplit("; e=", 1)[0] + f"; e={args.e}" if "; e=" in cfg["desc"] else cfg["desc"]
    if "; e=" not in cfg["desc"]:
        cfg["desc"] += f"; e={cfg['e']}"
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    assert swm.device_count() > 0, "bench needs a HIP device"
    forms = [f for f in args.forms.split(",") if f]
    assert forms and all(f in FORMS for f in forms), forms

    t0 = time.time()
    g = swm.synth_genome(cfg["n"], seed=GENOME_SEED, kind=cfg["kind"], frac=cfg["frac"])
    log(f"genome {cfg['n']} bp in {time.time() - t0:.1f}s")
    if args.copy_probe:
        a = swm.DeviceBuffer(256 << 20).zero()
        b = swm.DeviceBuffer(256 << 20)
        swm.lib().swm_copy_d2d(swm.C.c_void_p(b.ptr), swm.C.c_void_p(a.ptr), 256 << 20)
        a.free()
        b.free()
    if args.index_only:
        times = [swm.build_index(g, 12).build_ms for _ in range(args.index_only)]
        print(json.dumps({"index_build_ms": [round(t, 3) for t in times], "bases": cfg["n"]}), flush=True)
        return
    ix = swm.build_index(g, 12)
    ix2 = swm.build_index(g, 12)      # timed again with the device warm
    build_ms = min(ix.build_ms, ix2.build_ms)
    del ix2
    log(f"index: {ix.positions.size} positions, GPU build {build_ms:.1f} ms")
    # algorithmic bytes of the build: FASTA bytes + planes (N/4) + N mask (N/8)
    # + counts and offsets (2 x 4 B x 4^k) + positions and residuals (6 B each)
    build_bytes = cfg["n"] + cfg["n"] // 4 + cfg["n"] // 8 + 2 * 4 * 4 ** 12 + 6 * int(ix.positions.size)
    index_build = {"kernels": "k_encode, k_pack_hist, k_pack_pass0 (keys + sort pass 0), k_rs_pass x2 (the last "
                              "writes positions, residuals, starts), k_rs_cmax, k_rs_fix, k_counts, k_tail",
                   "ms": round(build_ms, 3),
                   "bound": "hbm", "algorithmic_bytes": build_bytes,
                   "achieved": round(build_bytes / (build_ms / 1e3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                   "unit": "GB/s", "frac": round(build_bytes / (build_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                   "bases": cfg["n"], "positions": int(ix.positions.size)}
    reads, lens, _, _ = swm.synth_reads(g, cfg["reads"], len_min=cfg["lmin"], len_max=cfg["lmax"],
                                        err_min=cfg["err"][0], err_max=cfg["err"][1], indel_share=0.2,
                                        seed=READS_SEED, stride=256)
    log(f"reads {cfg['reads']} in {time.time() - t0:.1f}s")
    m = swm.Mapper.from_index(ix)
    n = cfg["reads"]
    d_reads = swm.DeviceBuffer.from_array(reads)
    d_lens = swm.DeviceBuffer.from_array(lens)
    d_out = swm.DeviceBuffer(n * swm.RESULT_DTYPE.itemsize).zero()
    e = cfg["e"]
    for kv in args.set:
        name, val = kv.split("=", 1)
        opt = {v: k for k, v in swm.OPTION_NAMES.items()}[name]
        m.set_option(opt, int(val))
        log(f"option {name} = {val}")
    tuned = None
    if args.autotune:
        k = min(n, 100_000)
        tuned = m.autotune(reads[:k], lens[:k], e, reps=3)
        log(f"autotuned: {tuned}")

    def step():
        # swm_map_batch_device returns once the batch's results are in HBM
        # (it synchronizes the context stream), so each step is bracketed
        m.map_device(d_reads.ptr, 256, d_lens.ptr, n, e, d_out.ptr)
        return m.stats()

    results = {}
    ref_out = None
    for f in forms:
        elapsed, lat, st_all = time_form(m, FORMS[f], step, args.steps, args.warmup)
        out = d_out.to_array(swm.RESULT_DTYPE)
        if ref_out is None:
            ref_out = out
        same = bool(np.array_equal(out, ref_out))
        nv = st_all[-1]["n_verifications"]
        results[f] = {"elapsed": elapsed, "lat": lat, "st": st_all, "nv": nv, "same": same}
        log(f"form {f}: {elapsed * 1e3 / args.steps:.3f} ms/step, {nv * args.steps / elapsed / 1e6:.1f} M verif/s, "
            f"same results as {forms[0]}: {same}")
        for s in st_all:
            log("  " + " ".join(f"{k}={s[k]:.2f}" for k in ("ms_seed", "ms_plan", "ms_compact", "ms_myers",
                                                             "ms_emit", "ms_total"))
                + " " + " ".join(f"{k}={s[k]}" for k in ("n_heavy_items", "n_large_batches", "n_heavy_reads")))
    # per-read candidate histogram of the workload (one extra run, synced form)
    m.set_option(swm.OPT_GRAPH, 0)
    m.set_option(swm.OPT_HISTOGRAM, 1)
    hs = step()
    m.set_option(swm.OPT_HISTOGRAM, 0)
    hist = {("0" if b == 0 else f"{1 << (b - 1)}-{(1 << b) - 1}"): int(v) for b, v in enumerate(hs["cand_hist"]) if v}
    log(f"candidates per read: {hist}, max {hs['max_cands']}")
    records = None if args.no_records else records_bench(swm, m, reads, lens, e, ref_out, cfg["n"], log)
    best = min(results, key=lambda f: results[f]["elapsed"])
    R = results[best]
    nv = R["nv"]
    verif_per_s = nv * args.steps / R["elapsed"]
    reads_per_s = n * args.steps / R["elapsed"]

    timed = next((f for f in ("queue", "plain") if f in results), None)
    roofs = {"roofline": None}
    if timed:
        roofs = rooflines(swm, results[timed]["st"], lens, args.config, timed)

    serving = None
    if args.serve_seconds > 0:
        wq = args.serve_wq if args.serve_wq is not None else FORMS[best][swm.OPT_WORK_QUEUE]
        m.set_option(swm.OPT_WORK_QUEUE, wq)
        log(f"serving with work queues {'on' if wq else 'off'}")
        serving = serve_bench(swm, m, g, e, args.serve_seconds, log, slots=args.serve_slots,
                              graph=args.serve_form == "graph", merge=args.serve_merge)

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    cpu = None
    if not args.no_cpu:
        ref = orc.OracleRef(12, ix.hash_count, ix.positions, ix.pos_lseeds, ix.base_2bits)
        threads = len(os.sched_getaffinity(0))

        def timed_sample(nthreads, seconds, probe):
            tp = time.perf_counter()
            ref.map(reads[:probe], lens[:probe], e, nthreads)
            dtp = time.perf_counter() - tp
            ns = int(min(n, max(probe, probe * seconds / max(dtp, 1e-3))))
            tc = time.perf_counter()
            out, nv_c = ref.map(reads[:ns], lens[:ns], e, nthreads)
            return out, nv_c, ns, time.perf_counter() - tc

        cpu_out, nv_cpu, ns, dtc = timed_sample(threads, args.cpu_seconds, min(n, 4000))
        same = all(np.array_equal(cpu_out[f], ref_out[f][:ns]) for f in FIELDS)
        one_out, nv_one, ns1, dt1 = timed_sample(1, args.cpu_seconds / 3, min(n, 500))
        # the serial index build (serial_index.c) at the workload's size, one thread
        ti = time.perf_counter()
        oix = orc.OracleIndex(g, 12)
        dti = time.perf_counter() - ti
        ix_same = all(np.array_equal(a, ix.files()[k]) for k, a in oix.files().items())
        del oix
        cpu = {"value": round(nv_cpu / dtc, 1), "unit": "verifications/s", "cores": threads, "kind": "port",
               "reads_per_s": round(ns / dtc, 1), "gpu_results_identical": same,
               "sample": f"first {ns} reads of the same batch, oracle (C restatement, OpenMP dynamic,1) on every "
                         f"thread of the process affinity, {dtc:.1f}s",
               "host": host_cpu(),
               "single_thread": {"value": round(nv_one / dt1, 1), "unit": "verifications/s",
                                 "reads_per_s": round(ns1 / dt1, 1), "cores": 1,
                                 "sample": f"first {ns1} reads, 1 thread, {dt1:.1f}s"},
               "index_build": {"ms": round(dti * 1e3, 1), "threads": 1, "bases": cfg["n"],
                               "identical_to_gpu_build": ix_same,
                               "what": "oracle restatement of serial_index.c construct_hash_index, whole genome"}}
        log(f"cpu baseline: {cpu}")
    # end to end: reads from (pageable) host memory, results back to host
    for k_, v_ in FORMS[best].items():
        m.set_option(k_, v_)
    m.map(reads, lens, e)
    t_e = time.perf_counter()
    e2e_out = m.map(reads, lens, e)
    dt_e = time.perf_counter() - t_e
    end_to_end = {"reads_per_s": round(n / dt_e, 1), "ms": round(dt_e * 1e3, 3),
                  "verifications_per_s": round(nv / dt_e, 1),
                  "identical_results": bool(np.array_equal(e2e_out, ref_out)),
                  "entry": "swm_map_batch", "form": best,
                  "note": "reads copied in from pageable host memory and results copied back, inside the timing"}
    log(f"end to end: {end_to_end}")
    del e2e_out
    m.close()
    for b_ in (d_reads, d_lens, d_out):
        b_.free()
    configs = {}
    for name in [c for c in args.configs.split(",") if c]:
        if name == "small":
            c = CONFIGS["small"]
            gs = swm.synth_genome(c["n"], seed=GENOME_SEED, kind=c["kind"], frac=c["frac"])
            ixs = swm.build_index(gs, 12)
            configs["small"] = measure_config(swm, orc, "small", c, gs, ixs, ["queue", "graph"], args, log, c["e"],
                                              check_s=60.0)
            del gs, ixs
        elif name == "uneven":
            c = CONFIGS["uneven"]
            if (c["n"], c["kind"], c["frac"]) == (cfg["n"], cfg["kind"], cfg["frac"]):
                gu, ixu = g, ix                    # the headline's genome (the same seed and shape)
            else:
                gu = swm.synth_genome(c["n"], seed=GENOME_SEED, kind=c["kind"], frac=c["frac"])
                ixu = swm.build_index(gu, 12)
            configs["uneven"] = {f"e{ee}": measure_config(swm, orc, "uneven", c, gu, ixu, ["graph"], args, log, ee,
                                                          tune=True)
                                 for ee in (3, 5, 8)}
            del gu, ixu
        elif name == "skew":
            c = CONFIGS["skew"]
            gk = swm.synth_genome(c["n"], seed=GENOME_SEED, kind=c["kind"], frac=c["frac"])
            ixk = swm.build_index(gk, 12)
            configs["skew"] = measure_config(swm, orc, "skew", c, gk, ixk, ["plain", "queue"], args, log, c["e"],
                                             heavy=100)
            del gk, ixk
    lat_block = {"p50": round(float(np.percentile(R["lat"], 50)), 3),
                 "p99": round(float(np.percentile(R["lat"], 99)), 3), "of": f"whole 1-batch step, form {best}"}
    if serving:
        mid = serving["loads"][1]
        lat_block = {"p50": mid["p50_ms"], "p99": mid["p99_ms"],
                     "of": f"serving batches at {mid['offered']} offered load"}
    res = {
        "metric": "candidate verifications/sec + reads/sec at e=5; p50/p99 batch latency",
        "value": round(verif_per_s, 1), "unit": "verifications/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(R["elapsed"] * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "reads_per_s": round(reads_per_s, 1), "verifications_per_step": nv,
        "verifications_per_read": round(nv / n, 3),
        "candidates_per_read_hist": {"bins": hist, "max": int(hs["max_cands"])},
        "batch_latency_ms": lat_block,
        "forms": {f: {"ms_per_step": round(r["elapsed"] * 1e3 / args.steps, 3),
                      "verifications_per_s": round(r["nv"] * args.steps / r["elapsed"], 1),
                      "reads_per_s": round(n * args.steps / r["elapsed"], 1), "identical_results": r["same"]}
                  for f, r in results.items()},
        "config": {"workload": cfg["desc"], "name": args.config, "form": best, "reads": n, "e": e,
                   "tuned": tuned},
        "end_to_end": end_to_end,
        **roofs, "index_build": index_build,
        "serving": serving, "records": records, "cpu_baseline": cpu, "configs": configs,
        "child_processes_at_exit": len(child_processes()),
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
