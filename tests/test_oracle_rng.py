"""Oracle XORWOW against the reference's golden vectors (CPU only).

Golden data: tests/golden/xorwow.json, transcribed from
test/celeritas/random/XorwowRngEngine.test.cc:36-260 and
test/celeritas/random/RngReseed.test.cc:33-90.
"""

import ctypes
import json
import os

import numpy as np
import pytest

from tests import oracle_bind as ob

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "xorwow.json")))


def test_seeded_states_stream0():
    st = ob.rng_seeded_states(GOLDEN["seed"], 0, 8)
    np.testing.assert_array_equal(st.ravel(), GOLDEN["host_stream0_8slots_flattened"])


def test_seeded_states_stream1():
    st = ob.rng_seeded_states(GOLDEN["seed"], 1, 8)
    np.testing.assert_array_equal(st.ravel()[:8], GOLDEN["host_stream1_first8"])


def _canonical(words, n, kind):
    w = np.asarray(words, dtype=np.uint32)
    out = np.zeros(n, dtype=np.float64 if kind == 1 else np.float32)
    ob.load().orc_canonical_from_words.argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_int, ctypes.c_void_p]
    ob.load().orc_canonical_from_words(w.ctypes.data, n, kind, out.ctypes.data)
    return out


def test_canonical_double_bits():
    g = GOLDEN["canonical_double"]
    out = _canonical(g["words"], len(g["expected_hex"]), 1)
    expected = [float.fromhex(h) for h in g["expected_hex"]]
    assert list(out) == expected
    assert out[-1] < 1.0 and out[-1] == np.nextafter(1.0, 0.0)


def test_canonical_float_within_4ulp():
    # The reference compares with EXPECT_FLOAT_EQ (4 ULP); the largest words
    # round to 1.0f in single precision, which that comparison accepts.
    g = GOLDEN["canonical_float"]
    out = _canonical(g["words"], len(g["expected_hex"]), 2)
    expected = np.array([float.fromhex(h) for h in g["expected_hex"]], dtype=np.float32)
    ulp = np.abs(out.view(np.int32).astype(np.int64) - expected.view(np.int32).astype(np.int64))
    assert np.all(ulp <= 4)
    # Exact values the single-precision product must give
    assert out[0] == 0.0 and out[3] == 0.5 and out[-1] == 1.0


@pytest.mark.parametrize("offset", [0, 1, 2, 3, 4, 5, 15, 16, 17, 63, 255, 1000, 4096, 65535, 65536])
def test_jump_equals_repeated_next(offset):
    base = ob.rng_initialize(GOLDEN["seed"], 0, 0)
    draws, _ = ob.rng_draw(base, offset + 1)
    skipped = ob.rng_initialize(GOLDEN["seed"], 0, offset)
    first, _ = ob.rng_draw(skipped, 1)
    assert first[0] == draws[offset]


def test_discard_counts():
    s = ob.rng_initialize(GOLDEN["seed"], 0, 0)
    skip = s.copy()
    for count in (4, 21, 170, 65535):
        skip = ob.rng_discard(skip, count)
        draws, s = ob.rng_draw(s, count)
        a, s = ob.rng_draw(s, 1)
        b, skip = ob.rng_draw(skip, 1)
        assert a[0] == b[0]


def test_subsequence_is_2_pow_67():
    # (2^64 - 1 + 1) * 8 = 2^67 draws: the next subsequence (XorwowRngEngine.test.cc:240-258)
    sub = 1 << 19
    s = ob.rng_initialize(GOLDEN["seed"], sub, 0)
    for _ in range(8):
        s = ob.rng_discard(s, (1 << 64) - 1)
        s = ob.rng_discard(s, 1)
    s = ob.rng_discard(s, 1023)
    t = ob.rng_initialize(GOLDEN["seed"], sub + 1, 1023)
    a, _ = ob.rng_draw(s, 1)
    b, _ = ob.rng_draw(t, 1)
    assert a[0] == b[0]


def test_reseed_golden():
    g = GOLDEN["reseed"]
    n, ev = g["size"], g["event_id"]
    values = []
    for i in range(0, n, g["stride"]):
        s = ob.rng_initialize(GOLDEN["seed"], ev * n + i, 0)
        d, _ = ob.rng_draw(s, 1)
        values.append(int(d[0]))
    assert values == g["expected_first_draw"]


def test_canonical_moments():
    s = ob.rng_initialize(GOLDEN["seed"], 3, 0)
    x, _ = ob.rng_draw(s, 1 << 18, "f64")
    assert abs(x.mean() - 0.5) < 3e-3
    assert abs(x.var() - 1 / 12) < 2e-3
    assert x.min() >= 0 and x.max() < 1
