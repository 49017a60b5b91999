"""The read-encoding recipe the kernels apply (swm_recipe.c) reproduces the
oracle's literal encoder for every read length 1..254 and several seed
lengths (CPU)."""
import ctypes as C

import numpy as np
import pytest

import swmapper_amd as swm
import oracle_lib as orc


@pytest.mark.parametrize("k", [8, 10, 12, 13])
def test_recipe_matches_literal_encoder(k):
    L_fn = swm.lib().swm_encode_read_host
    L_fn.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]
    L_fn.restype = C.c_int
    rng = np.random.default_rng(k)
    for L in range(1, min(255, 16 * (k + 3))):
        for _ in range(3):
            r = np.frombuffer(b"ACGTN", np.uint8)[rng.choice(5, L, p=[.24, .24, .24, .24, .04])].copy()
            buf = np.zeros(L + 64, np.uint8)
            buf[:L] = r
            got = np.zeros(32, np.uint32)
            assert L_fn(buf.ctypes.data, L, k, got.ctypes.data) == 0
            exp, _ = orc.encode_read(r.tobytes(), k)
            assert np.array_equal(got, exp), (L, k)
