"""GPU parity of every pipeline form and tunable: the plain one-thread-per-item
kernels, the device work queues, the hipGraph-replayed fixed-capacity form
(including its overflow fallback) and the autotuner's choices all give the
oracle's results bit for bit.

The tandem-repeat genome gives reads with up to ~2600 records, so the
output-block flush (mapper_cpe.c:1790, 1858) and the duplicate rule after a
flush (:475-477) run in the cooperative emission kernel.
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
        assert bad.size == 0, f"{what} {f}: {bad.size} reads differ, first {bad[:5]} got {got[bad[:3]]} exp {exp[bad[:3]]}"


@pytest.fixture(scope="module")
def tandem():
    g = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    r, l, _, _ = wl.reads(g, 3000, seed=42)
    gi = swm.build_index(g, 12)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    exp = {e: ref.map(r, l, e) for e in (5, 8)}
    assert exp[8][0]["count"].max() > 600          # several output blocks per strand
    assert (l == 150).all()
    return gi, r, l, exp


SETTINGS = [
    {},
    {swm.OPT_WORK_QUEUE: 0},
    {swm.OPT_HEAVY_CANDS: 1},                      # every read through the cooperative kernels
    {swm.OPT_HEAVY_CANDS: 65535},
    {swm.OPT_MYERS_BLOCK: 64},
    {swm.OPT_MYERS_BLOCK: 128, swm.OPT_PLAN_HEAVY_WAVES: 1, swm.OPT_EMIT_HEAVY_WAVES: 1},
    {swm.OPT_GRAPH: 1},
    {swm.OPT_GRAPH: 1, swm.OPT_WORK_QUEUE: 0},
    {swm.OPT_MYERS_PACKED: 0},
    {swm.OPT_MYERS_PACKED: 1, swm.OPT_MYERS_BLOCK: 64},
    {swm.OPT_PLAN_OVERLAP: 1},                     # heavy plan kernels on a second stream beside the light-item one
    {swm.OPT_PLAN_LIGHT_WAVES: 64, swm.OPT_PLAN_HEAVY_WAVES: 8},
    {swm.OPT_GRAPH: 1, swm.OPT_PLAN_OVERLAP: 1, swm.OPT_PLAN_LIGHT_WAVES: 8},   # the fork and join inside a captured graph
]


@pytest.mark.parametrize("e", [5, 8])
@pytest.mark.parametrize("si", range(len(SETTINGS)))
def test_forms_match_oracle(tandem, e, si):
    gi, r, l, exp = tandem
    opts = SETTINGS[si]
    _say("forms tandem e", e, {swm.OPTION_NAMES[k]: v for k, v in opts.items()})
    m = swm.Mapper.from_index(gi)
    for k, v in opts.items():
        m.set_option(k, v)
    ref, nv = exp[e]
    # graph form: the first call runs synced and sizes the capacities, the
    # later ones replay the captured graph
    for rep in range(3 if opts.get(swm.OPT_GRAPH) else 1):
        got = m.map(r, l, e)
        _assert_same(got, ref, f"rep {rep}")
        st = m.stats()
        assert st["n_verifications"] == nv
        # every read has 150 bases: one band step per base per verification
        assert st["n_band_steps"] == 150 * nv, (st["n_band_steps"], nv)
        assert m.stats()["n_records"] == int(ref["count"].sum())
    m.close()


def test_graph_overflow_falls_back():
    """A graph captured for light batches meets a heavy batch of the same
    size: the device stands down, the batch runs synced, results exact."""
    gu = wl.genome(400_000, seed=71)
    gt = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    g = np.concatenate([gu, gt])
    ri, li, _, _ = wl.reads(gu, 2000, seed=72)
    rh, lh, _, _ = wl.reads(gt, 2000, seed=73)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    m.set_option(swm.OPT_GRAPH, 1)
    for tag, (r, l) in [("light", (ri, li)), ("light", (ri, li)), ("heavy", (rh, lh)), ("heavy", (rh, lh)),
                        ("light", (ri, li))]:
        _say("graph overflow", tag)
        exp, nv = ref.map(r, l, 5)
        got = m.map(r, l, 5)
        _assert_same(got, exp, tag)
        assert m.stats()["n_verifications"] == nv
    m.close()


def test_graph_replays_between_other_work():
    """Replays of one captured graph interleaved with eager runs of other
    batches on the same context all match the oracle."""
    g = wl.genome(300_000, seed=81, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.2)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    m = swm.Mapper.from_index(swm.build_index(g, 12))
    a = wl.reads(g, 3000, len_min=50, len_max=250, err=(0.01, 0.04), seed=82)[:2]
    b = wl.reads(g, 1700, seed=83)[:2]
    ea, _ = ref.map(*a, 5)
    eb, _ = ref.map(*b, 5)
    for i in range(4):
        _say("graph replay", i)
        m.set_option(swm.OPT_GRAPH, 1)
        _assert_same(m.map(*a, 5), ea, f"graph {i}")
        m.set_option(swm.OPT_GRAPH, 0)
        _assert_same(m.map(*b, 5), eb, f"eager {i}")
    m.close()


def test_tandem_runs_the_large_batch_sort(tandem):
    """The oracle comparisons above cover the large-batch sort (batches of
    65 .. 2048 survivors, sorted and de-duplicated in registers, one wave
    per batch) only if the fixture produces such batches, with both heavy
    items (the seed heap) and light ones."""
    gi, r, l, exp = tandem
    _say("tandem large batches")
    m = swm.Mapper.from_index(gi)
    got = m.map(r, l, 5)
    _assert_same(got, exp[5][0], "large batches")
    st = m.stats()
    _say("stats", {k: st[k] for k in ("n_large_batches", "n_heavy_items", "n_heavy_reads")})
    assert st["n_large_batches"] > 200 and st["n_heavy_items"] > 100, st
    m.close()


def test_autotune_keeps_results(tandem):
    gi, r, l, exp = tandem
    _say("autotune")
    m = swm.Mapper.from_index(gi)
    chosen = m.autotune(r[:1500], l[:1500], 5, reps=1)
    assert chosen["myers_block"] in (64, 128, 256)
    assert m.options() == chosen
    got = m.map(r, l, 5)
    _assert_same(got, exp[5][0], "autotuned")
    m.close()


def test_option_ranges(tandem):
    """Every option's documented range: values outside it are refused
    (SWM_EINVAL) and leave the option as it was; the defaults read back."""
    gi, _, _, _ = tandem
    m = swm.Mapper.from_index(gi)
    defaults = m.options()
    assert defaults["plan_overlap"] == 0 and defaults["plan_light_waves"] == 32, defaults
    bad = {swm.OPT_PLAN_OVERLAP: [2, -1], swm.OPT_PLAN_LIGHT_WAVES: [0, 7, 65],
           swm.OPT_PLAN_HEAVY_WAVES: [0, 33], swm.OPT_SEED_GENERIC: [2]}
    for opt, vals in bad.items():
        for v in vals:
            with pytest.raises(swm.SwmError) as ex:
                m.set_option(opt, v)
            assert ex.value.rc == -1, (swm.OPTION_NAMES[opt], v)
    assert m.options() == defaults
    for opt, v in ((swm.OPT_PLAN_OVERLAP, 1), (swm.OPT_PLAN_LIGHT_WAVES, 8), (swm.OPT_PLAN_LIGHT_WAVES, 64)):
        m.set_option(opt, v)
        assert m.get_option(opt) == v
    m.close()
