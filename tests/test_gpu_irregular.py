"""Buckets out of ascending order.  The reference's serial scatter leaves
a few buckets unsorted (Q23: the zeroed tail starts overwrite the head of a
later bucket, their own bucket keeps a 0 slot); block_heap_sort
(mapper_cpe.c:1999-2089) then merges non-ascending blocks, and its output
order and surviving duplicates follow the heap's steps.  Here an index of a
tandem-repeat genome is perturbed on purpose -- buckets reversed, rotated,
given a zero tail -- so that most items select such a bucket, and the GPU
(which plans those items with k_plan_irreg) must give the oracle's results
and records in every form."""
import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc
import workloads as wl

pytestmark = pytest.mark.gpu

K = 12


def _say(*a):
    print("[case]", *a, flush=True)


def _same(got, exp, what):
    for f in ("count", "pos", "strand", "dist"):
        bad = np.flatnonzero(got[f] != exp[f])
        assert bad.size == 0, f"{what} {f}: {bad.size} reads differ, first {bad[:5]}"


@pytest.fixture(scope="module")
def world():
    g = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    oix = orc.OracleIndex(g, K)
    hc = oix.hash_count
    pos, ls = oix.positions.copy(), oix.pos_lseeds.copy()
    off = np.zeros(hc.size + 1, np.int64)
    np.cumsum(hc, out=off[1:])
    nperturbed = 0
    for c in np.flatnonzero(hc >= 2):
        a, b = int(off[c]), int(off[c + 1])
        kind = c % 5
        if kind == 0:                          # reversed
            pos[a:b], ls[a:b] = pos[a:b][::-1].copy(), ls[a:b][::-1].copy()
        elif kind == 1:                        # last entry moved to the head
            pos[a:b], ls[a:b] = np.roll(pos[a:b], 1), np.roll(ls[a:b], 1)
        elif kind == 2:                        # an unwritten (0) last slot
            pos[b - 1] = 0
        else:
            continue
        nperturbed += 1
    assert nperturbed > 500
    ix = swm.Index(K, oix.n_bases, oix.base_2bits, oix.base_3bits, hc, oix.hash_offset_end, pos, ls)
    ref = orc.OracleRef(K, hc, pos, ls, oix.base_2bits)
    r, l, _, _ = wl.reads(g, 2000, seed=43)
    return ix, ref, r, l


@pytest.mark.parametrize("e", [5, 8])
def test_irregular_buckets_every_form(world, e):
    ix, ref, r, l = world
    exp, nv = ref.map(r, l, e)
    m = swm.Mapper.from_index(ix)
    for form, opts in (("queue", {}), ("plain", {swm.OPT_WORK_QUEUE: 0}), ("graph", {swm.OPT_GRAPH: 1})):
        for k, v in opts.items():
            m.set_option(k, v)
        for rep in range(2 if form == "graph" else 1):
            _say("irregular", form, "e", e, "rep", rep)
            _same(m.map(r, l, e), exp, f"{form} e={e}")
            st = m.stats()
            assert st["n_verifications"] == nv
            assert st["n_irregular_items"] > 100, st["n_irregular_items"]
        for k in opts:
            m.set_option(k, 0)
    m.close()


def test_irregular_buckets_records(world):
    ix, ref, r, l = world
    m = swm.Mapper.from_index(ix)
    _say("irregular records")
    out, recs, text = m.map_records(r, l, 5)
    _, orecs, _ = ref.trace(r, l, 5)
    assert recs.size == orecs.size
    assert np.array_equal(recs["read"], orecs["read"]) and np.array_equal(recs["strand"], orecs["strand"])
    assert np.array_equal(recs["pos"], orecs["loc"]) and np.array_equal(recs["primary"], orecs["primary"])
    assert swm.cigar_strings(recs, text) == [bytes(c) for c in orecs["cigar"]]
    m.close()
