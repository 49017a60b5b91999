"""Serving loop on the GPU: continuous batching over the resident index.

Batches of mixed sizes and read lengths are submitted together to a server
with several slots (streams), so their work overlaps on the device; every
result must equal the same batch mapped alone, and the oracle.
"""
import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc
import workloads as wl

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _assert_same(got, exp, what=""):
    for f in ("count", "pos", "strand", "dist"):
        bad = np.flatnonzero(got[f] != exp[f])
        assert bad.size == 0, f"{what} {f}: {bad.size} reads differ, first {bad[:5]}"


@pytest.fixture(scope="module")
def setup():
    g = wl.genome(600_000, seed=91, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.1)
    pool_r, pool_l, _, _ = wl.reads(g, 30_000, len_min=50, len_max=250, err=(0.01, 0.04), seed=92)
    gi = swm.build_index(g, 12)
    rng = np.random.default_rng(93)
    batches = []
    for _ in range(14):
        n = int(rng.integers(200, 4000))
        s = int(rng.integers(0, pool_l.size - n))
        batches.append((pool_r[s:s + n], pool_l[s:s + n]))
    return g, gi, batches


@pytest.mark.parametrize("graph,slots,merge", [(True, 4, 0), (False, 3, 0), (True, 1, 0), (True, 3, 20000),
                                               (False, 2, 16384), (True, 1, 262144)])
def test_mix_equals_alone(setup, graph, slots, merge):
    g, gi, batches = setup
    m = swm.Mapper.from_index(gi)
    alone = [m.map(r, l, 5) for r, l in batches]
    _say("serve mix graph", graph, "slots", slots, "merge", merge)
    srv = swm.Server(m, e=5, stride=256, n_slots=slots, graph=graph, max_batch=4000, max_launch=merge)
    tickets = [srv.submit(r, l) for r, l in batches]
    infos = []
    for t, a in zip(tickets, alone):
        got, info = srv.wait(t)
        _assert_same(got, a, f"ticket {t}")
        assert info["status"] == 0 and info["t_submit"] <= info["t_admit"] <= info["t_done"]
        infos.append(info)
    tot = srv.totals()
    assert tot["batches"] == len(batches)
    assert tot["reads"] == sum(l.size for _, l in batches)
    assert tot["peak_bytes"] <= tot["budget"]
    for i in infos:
        assert i["launch_reads"] >= i["n_reads"] and (merge or i["launch_batches"] == 1)
        assert i["launch_reads"] <= max(merge, 4000)
    if merge and slots == 1:
        assert max(i["launch_batches"] for i in infos) > 1      # submitted together: merged
    srv.close()
    m.close()


def test_serve_matches_oracle(setup):
    g, gi, batches = setup
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    m = swm.Mapper.from_index(gi)
    srv = swm.Server(m, e=5, n_slots=4, graph=True, max_launch=0)
    _say("serve vs oracle")
    sel = batches[:5]
    tickets = [srv.submit(r, l) for r, l in sel]
    for t, (r, l) in zip(tickets, sel):
        got, info = srv.wait(t)
        exp, nv = ref.map(r, l, 5)
        _assert_same(got, exp, f"ticket {t}")
        assert info["stats"]["n_verifications"] == nv
    srv.close()
    m.close()


def test_admission_by_budget(setup):
    """A budget that holds one batch at a time still completes every batch in
    order of submission; a batch larger than the budget is refused."""
    g, gi, batches = setup
    m = swm.Mapper.from_index(gi)
    big = max(l.size for _, l in batches)
    srv = swm.Server(m, e=5, n_slots=4, graph=False, hbm_budget=int(big * 8000 + (1 << 21)))
    _say("serve budget")
    tickets = [srv.submit(r, l) for r, l in batches[:6]]
    admits = []
    for t, (r, l) in zip(tickets, batches[:6]):
        got, info = srv.wait(t)
        _assert_same(got, m.map(r, l, 5), f"ticket {t}")
        admits.append(info["t_admit"])
    assert admits == sorted(admits)
    tot = srv.totals()
    assert tot["peak_bytes"] <= tot["budget"]
    r, l = batches[0]
    reps = int(np.ceil(tot["budget"] / 4000 / l.size)) + 2
    huge_r = np.tile(r, (reps, 1))
    huge_l = np.tile(l, reps)
    if huge_l.size <= 65536:
        with pytest.raises(swm.SwmError):
            srv.submit(huge_r, huge_l)
    srv.close()
    m.close()


def test_empty_and_invalid_batches(setup):
    g, gi, batches = setup
    m = swm.Mapper.from_index(gi)
    srv = swm.Server(m, e=5, n_slots=2)
    _say("serve empty/invalid")
    t = srv.submit(np.zeros((0, 256), np.uint8), np.zeros(0, np.uint32))
    got, info = srv.wait(t)
    assert got.size == 0 and info["status"] == 0
    bad = np.full(3, 255, np.uint32)      # L + 2 > stride
    with pytest.raises(swm.SwmError):
        srv.submit(np.zeros((3, 256), np.uint8), bad)
    srv.close()
    m.close()


def test_no_capture_after_warmup(setup):
    """After warmup every launch size (powers of two from 1024 to 2^18 reads)
    replays a graph captured up front: traffic of mixed sizes, merged or not,
    captures nothing on the worker threads."""
    # a uniform genome: per-read totals vary little from batch to batch, so
    # no batch overflows the capacities learned at warmup (an overflow would
    # rightly recapture)
    g = wl.genome(600_000, seed=95)
    pool_r, pool_l, _, _ = wl.reads(g, 30_000, len_min=50, len_max=250, err=(0.01, 0.04), seed=96)
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    srv = swm.Server(m, e=5, stride=256, n_slots=2, graph=True, max_batch=65536, max_launch=262144)
    _say("serve warmup")
    srv.warmup(pool_r, pool_l)
    rng = np.random.default_rng(94)
    work = []
    for n in (1000, 1500, 2100, 3000, 5000, 9000, 17000, 33000, 65000, 40000, 1024, 64000):
        idx = rng.integers(0, pool_l.size, n)
        work.append((pool_r[idx], pool_l[idx]))
    _say("serve mixed sizes")
    tickets = [srv.submit(r, l) for r, l in work]
    for t, (r, l) in zip(tickets, work):
        got, info = srv.wait(t)
        assert info["stats"]["n_graph_captures"] == 0, info["launch_reads"]
        _assert_same(got, m.map(r, l, 5), f"ticket {t}")
    srv.close()
    m.close()
