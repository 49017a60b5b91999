"""The oracle's hand-worked known answers (tests/test_oracle_pins.py) run
through the GPU path: the fabricated seed-heap index gives the same number of
verifications (1087 in the first batch, 63 in the second, seven of them
starts verified again) in every form, and the repeat indels give the same
records and CIGAR text."""
import numpy as np
import pytest

import swmapper_amd as swm
import test_oracle_pins as pins

pytestmark = pytest.mark.gpu

FORMS = {"plain": {swm.OPT_WORK_QUEUE: 0}, "queue": {}, "graph": {swm.OPT_GRAPH: 1},
         "giant lists": {swm.OPT_GIANT_POS: 100, swm.OPT_GIANT_CANDS: 100}}


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("form", sorted(FORMS))
def test_seed_heap_world_verifications(form):
    hc, pos, ls, b2 = pins.heap_world()
    ix = swm.Index(pins.K, 20480, b2, np.zeros(640, np.uint32), hc, np.zeros(hc.size + 1, np.uint32), pos, ls)
    m = swm.Mapper.from_index(ix)
    for k, v in FORMS[form].items():
        m.set_option(k, v)
    reads = np.zeros((4, 32), np.uint8)
    reads[:, :15] = np.frombuffer(pins.READ, np.uint8)
    lens = np.full(4, 15, np.uint32)
    for rep in range(3 if form == "graph" else 1):
        _say("heap world", form, rep)
        out = m.map(reads, lens, 0)
        st = m.stats()
        assert (out["count"] == 0).all()
        assert st["n_verifications"] == 4 * (len(pins.HEAP_BATCH0) + len(pins.HEAP_BATCH1)), st["n_verifications"]
        assert st["n_band_steps"] == 15 * st["n_verifications"]
    if form == "giant lists":
        assert st["n_giant_items"] == 4, st
    m.close()


def test_traceback_world_records():
    g, g2 = pins.tb_world()
    cases = pins.tb_cases(g, g2)
    for gen in (g, g2):
        m = swm.Mapper.from_index(swm.build_index(gen, pins.K))
        names = [n for n, (gg, _, _) in cases.items() if gg is gen]
        reads = np.zeros((len(names), 256), np.uint8)
        lens = np.zeros(len(names), np.uint32)
        for i, n in enumerate(names):
            r = cases[n][1]
            reads[i, :len(r)] = np.frombuffer(r, np.uint8)
            lens[i] = len(r)
        _say("traceback world", names)
        out, recs, text = m.map_records(reads, lens, 5)
        cig = swm.cigar_strings(recs, text)
        for i, n in enumerate(names):
            sel = np.flatnonzero(recs["read"] == i)
            got = [(int(recs["pos"][j]), cig[j], int(recs["dist"][j]), int(recs["strand"][j])) for j in sel]
            dist = 2 if "dinucleotide" in n else 1
            assert got == [(pins.LOC + 1, cases[n][2], dist, 0)], (n, got)
            assert out["pos"][i] == pins.LOC + 1 and out["count"][i] == 1
        m.close()
