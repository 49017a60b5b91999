"""Oracle parity at the benchmark configurations' full sizes, on the very
inputs bench.py times (swmapper_amd.configs: same genomes, reads and seeds).

  small    4 Mbp uniform, 100k x 150 bp, e = 5: every read
  repeat5  256 Mbp + 5% repeat families: the GPU index build byte for byte
           against the oracle's serial build; 1M x 150 bp at e = 5, every read
           (graph form) and the plain form against it
  uneven   the same genome, 1M reads of 50-250 bp at e = 3, 5, 8 (graph
           form; e = 5 after the autotuner), every read
  skew     256 Mbp + 30% tandem repeats: index byte for byte; 500k x 150 bp
           at e = 5 with per-read candidate counts up to ~10^5: the work-queue
           and plain forms agree on every read, the giant lists (items over
           16384 positions, reads over 16384 candidates) run, and the oracle
           agrees on the 2000 reads with the most records plus a uniform
           sample of 20000

The oracle maps with every host core (OpenMP)."""
import os

import numpy as np
import pytest

import swmapper_amd as swm
from swmapper_amd import configs as cf
import oracle_lib as orc

pytestmark = pytest.mark.gpu

FIELDS = ("count", "pos", "strand", "dist")


def _say(*a):
    print("[case]", *a, flush=True)


def _same(got, exp, what):
    for f in FIELDS:
        bad = np.flatnonzero(got[f] != exp[f])
        assert bad.size == 0, f"{what} {f}: {bad.size} reads differ, first {bad[:5]}"


def _threads():
    return len(os.sched_getaffinity(0))


class World:
    def __init__(self, name):
        _say("world", name)
        self.name = name
        self.g = cf.genome(name)
        self.ix = swm.build_index(self.g, 12)
        self.oix = orc.OracleIndex(self.g, 12)
        self.ref = orc.OracleRef.from_index(self.oix)
        self.m = swm.Mapper.from_index(self.ix)


@pytest.fixture(scope="module")
def repeat5():
    w = World("repeat5")
    yield w
    w.m.close()


@pytest.fixture(scope="module")
def skew():
    w = World("skew")
    yield w
    w.m.close()


def _index_identical(w):
    for ext, arr in w.oix.files().items():
        got = w.ix.files()[ext]
        assert got.dtype == arr.dtype and got.shape == arr.shape, ext
        bad = np.flatnonzero(got != arr)
        assert bad.size == 0, f"{w.name} {ext}: {bad.size} words differ, first {bad[:5]}"


@pytest.mark.timeout(300)
def test_small_every_read():
    g = cf.genome("small")
    r, l = cf.reads("small", g)
    ix = swm.build_index(g, 12)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    exp, nv = ref.map(r, l, 5, _threads())
    m = swm.Mapper.from_index(ix)
    for form, opts in (("plain", {swm.OPT_WORK_QUEUE: 0}), ("graph", {swm.OPT_GRAPH: 1})):
        _say("small", form)
        for k, v in opts.items():
            m.set_option(k, v)
        for _ in range(2 if form == "graph" else 1):
            _same(m.map(r, l, 5), exp, f"small {form}")
            assert m.stats()["n_verifications"] == nv
    m.close()


@pytest.mark.timeout(300)
def test_repeat5_index_256mbp_identical(repeat5):
    _say("repeat5 index 256 Mbp")
    _index_identical(repeat5)


@pytest.mark.timeout(300)
def test_repeat5_every_read(repeat5):
    w = repeat5
    r, l = cf.reads("repeat5", w.g)
    exp, nv = w.ref.map(r, l, 5, _threads())
    for form, opts in (("graph", {swm.OPT_GRAPH: 1, swm.OPT_WORK_QUEUE: 1}),
                       ("plain", {swm.OPT_GRAPH: 0, swm.OPT_WORK_QUEUE: 0})):
        _say("repeat5 1M", form)
        for k, v in opts.items():
            w.m.set_option(k, v)
        for _ in range(2 if form == "graph" else 1):
            _same(w.m.map(r, l, 5), exp, f"repeat5 {form}")
            assert w.m.stats()["n_verifications"] == nv
    w.m.set_option(swm.OPT_WORK_QUEUE, 1)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("e", [3, 5, 8])
def test_uneven_every_read(repeat5, e):
    w = repeat5
    r, l = cf.reads("uneven", w.g)
    assert l.min() == 50 and l.max() == 250
    exp, nv = w.ref.map(r, l, e, _threads())
    w.m.set_option(swm.OPT_GRAPH, 1)
    if e == 5:
        _say("uneven autotune")
        w.m.autotune(r[:100_000], l[:100_000], e, reps=1)
        w.m.set_option(swm.OPT_GRAPH, 1)
    for rep in range(2):
        _say("uneven 1M e", e, "rep", rep)
        _same(w.m.map(r, l, e), exp, f"uneven e={e}")
        assert w.m.stats()["n_verifications"] == nv


@pytest.mark.timeout(300)
def test_skew_index_256mbp_identical(skew):
    _say("skew index 256 Mbp")
    _index_identical(skew)


@pytest.mark.timeout(300)
def test_skew_forms_agree_and_giant_lists_run(skew):
    w = skew
    r, l = cf.reads("skew", w.g)
    out = {}
    for form, opts in (("queue", {swm.OPT_WORK_QUEUE: 1, swm.OPT_GRAPH: 0}),
                       ("graph", {swm.OPT_WORK_QUEUE: 1, swm.OPT_GRAPH: 1}),
                       ("plain", {swm.OPT_WORK_QUEUE: 0, swm.OPT_GRAPH: 0})):
        _say("skew 500k", form)
        for k, v in opts.items():
            w.m.set_option(k, v)
        for _ in range(2 if form == "graph" else 1):
            out[form] = w.m.map(r, l, 5)
            st = w.m.stats()
        if form == "queue":
            _say("skew stats", {k: st[k] for k in ("n_giant_items", "n_giant_reads", "n_heavy_items",
                                                    "n_heavy_reads", "n_verifications")})
            assert st["n_giant_items"] > 0 and st["n_giant_reads"] > 0, st
        out[form + "_nv"] = st["n_verifications"]
    _same(out["graph"], out["queue"], "skew graph vs queue")
    _same(out["plain"], out["queue"], "skew plain vs queue")
    assert out["plain_nv"] == out["queue_nv"] == out["graph_nv"]
    w.m.set_option(swm.OPT_WORK_QUEUE, 1)
    w.m.set_option(swm.OPT_GRAPH, 0)
    w.queue_out = out["queue"]


@pytest.mark.timeout(300)
def test_skew_oracle_heaviest_and_sample(skew):
    w = skew
    r, l = cf.reads("skew", w.g)
    got = getattr(w, "queue_out", None)
    if got is None:
        got = w.m.map(r, l, 5)
    heavy = np.argsort(got["count"], kind="stable")[::-1][:2000]
    rng = np.random.default_rng(5)
    sample = np.sort(rng.choice(r.shape[0], 20_000, replace=False))
    for tag, idx in (("heaviest 2000", np.sort(heavy)), ("uniform 20000", sample)):
        _say("skew oracle", tag, "max records", int(got["count"][idx].max()))
        exp, _ = w.ref.map(r[idx], l[idx], 5, _threads())
        _same(got[idx], exp, f"skew {tag}")
    # the oracle's check covers reads past the giant-read threshold
    assert got["count"][heavy].max() > 16384 and got["count"][heavy].min() > 100


@pytest.mark.timeout(300)
def test_giant_lists_in_every_form_on_the_tandem_fixture():
    """The tandem fixture with the giant thresholds lowered: giant plan items
    and giant emission reads in the work-queue and graph forms, results equal
    to the oracle's."""
    import workloads as wl
    g = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    r, l, _, _ = wl.reads(g, 3000, seed=42)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    for e in (5, 8):
        exp, nv = ref.map(r, l, e)
        for form, opts in (("queue", {swm.OPT_GRAPH: 0}), ("graph", {swm.OPT_GRAPH: 1})):
            for gp, gc in ((256, 64), (1, 1), (4096, 1024)):
                _say("giant lists", form, "e", e, "thresholds", gp, gc)
                m.set_option(swm.OPT_GIANT_POS, gp)
                m.set_option(swm.OPT_GIANT_CANDS, gc)
                for k, v in opts.items():
                    m.set_option(k, v)
                for _ in range(2 if form == "graph" else 1):
                    _same(m.map(r, l, e), exp, f"giant {form} e={e} {gp}/{gc}")
                st = m.stats()
                assert st["n_verifications"] == nv
                if form == "queue":
                    assert st["n_giant_items"] > 0 and st["n_giant_reads"] > 0, st
    m.close()
