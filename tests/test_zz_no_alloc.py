"""The no-allocation entry point (swm_map_batch_device_noalloc) against
swm_map_batch_device, the entry point it mirrors, through the C ABI: results
identical read for read on two calls in a row (the second smaller, so that
it runs over what the first left in the scratch), in the work-queue and the
plain form; an oversized batch is refused with SWM_ENOSPC and the next call
is exact again."""
import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc
import workloads as wl

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _same(got, exp, what):
    for f in ("count", "pos", "strand", "dist"):
        bad = np.flatnonzero(got[f] != exp[f])
        assert bad.size == 0, f"{what} {f}: {bad.size} reads differ, first {bad[:5]}"


def _run(m, fn, r, l, e):
    n, stride = r.shape
    dr, dl = swm.DeviceBuffer.from_array(r), swm.DeviceBuffer.from_array(l)
    do = swm.DeviceBuffer(n * swm.RESULT_DTYPE.itemsize)
    try:
        getattr(m, fn)(dr.ptr, stride, dl.ptr, n, e, do.ptr)
        return do.to_array(swm.RESULT_DTYPE)
    finally:
        for b in (dr, dl, do):
            b.free()


@pytest.fixture(scope="module")
def genome():
    g = wl.genome(1_000_000, seed=91, kind=swm.GENOME_REPEAT_FAMILIES, frac=0.05)
    return g, swm.build_index(g, 12)


@pytest.mark.parametrize("work_queue", [1, 0])
def test_noalloc_matches_device_entry(genome, work_queue):
    g, gi = genome
    m = swm.Mapper.from_index(gi)
    m.set_option(swm.OPT_WORK_QUEUE, work_queue)
    ref = orc.OracleRef.from_index(orc.OracleIndex(g, 12))
    big = wl.reads(g, swm.NOALLOC_MAX_READS, len_min=50, len_max=250, err=(0.01, 0.04), seed=92)[:2]
    small = wl.reads(g, 2500, seed=93)[:2]
    for tag, (r, l), e in (("8192 reads L 50-250", big, 5), ("2500 reads L 150", small, 3)):
        _say("noalloc", tag, "e", e, "work_queue", work_queue)
        exp = _run(m, "map_device", r, l, e)
        nv = m.stats()["n_verifications"]
        got = _run(m, "map_device_noalloc", r, l, e)
        _same(got, exp, tag)
        assert m.stats()["n_verifications"] == nv
        _same(got, ref.map(r, l, e)[0], tag + " vs oracle")
    m.close()


def test_noalloc_refuses_what_does_not_fit(genome):
    g, gi = genome
    m = swm.Mapper.from_index(gi)
    r, l = wl.reads(g, swm.NOALLOC_MAX_READS + 1, seed=94)[:2]
    _say("noalloc refuses", r.shape[0], "reads")
    with pytest.raises(swm.SwmError) as ei:
        _run(m, "map_device_noalloc", r, l, 5)
    assert ei.value.rc == swm.SWM_ENOSPC
    # a batch within the read limit whose candidates exceed the scratch: the
    # device stands down at the capacity check, the call says so
    gt = wl.genome(400_000, seed=41, kind=swm.GENOME_TANDEM, frac=0.3)
    mt = swm.Mapper.from_index(swm.build_index(gt, 12))
    rt, lt = wl.reads(gt, swm.NOALLOC_MAX_READS, seed=95)[:2]
    _say("noalloc tandem", rt.shape[0], "reads")
    with pytest.raises(swm.SwmError) as ei:
        _run(mt, "map_device_noalloc", rt, lt, 8)
    assert ei.value.rc == swm.SWM_ENOSPC
    # and the next call, smaller, is exact again
    rs, ls = rt[:300], lt[:300]
    _say("noalloc tandem after refusal", rs.shape[0], "reads")
    _same(_run(mt, "map_device_noalloc", rs, ls, 8), _run(mt, "map_device", rs, ls, 8), "after refusal")
    mt.close()
    m.close()
