"""The measurement probes bench.py quotes rooflines against: the int32 VALU
rate (`swm_valu_probe`) and the random 8-byte gather rate
(`swm_gather_probe`).  Checked for sane magnitudes on an MI355X (256 CUs at
up to 2.4 GHz, HBM3E at 8 TB/s) and for their argument checks.
"""
import ctypes as C

import pytest

import swmapper_amd as swm

pytestmark = pytest.mark.gpu


def test_valu_probe_magnitude():
    ops = swm.valu_peak()
    # 256 CUs x 4 SIMDs x 16 lanes x ~2.1-2.4 GHz: 34-40 T lane ops/s
    assert 20e12 < ops < 60e12, ops


def test_gather_probe_magnitude():
    rate = swm.gather_peak(64 << 20)
    # one 128-byte line per gather: below the HBM peak in lines (8 TB/s / 128 B = 62.5 G/s)
    # and far above a serial rate
    assert 5e9 < rate < 80e9, rate


def test_gather_probe_rejects_bad_arguments():
    v = C.c_double(0)
    assert swm.lib().swm_gather_probe(1 << 10, C.byref(v)) == -1      # under 1 MiB
    assert swm.lib().swm_gather_probe(64 << 20, None) == -1
