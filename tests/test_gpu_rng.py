"""HIP XORWOW kernels against the oracle and the reference golden words."""

import json
import os

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import problems

pytestmark = pytest.mark.gpu
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "xorwow.json")))


def rng_problem(seed):
    p = problems.two_boxes_track_init()
    p["seed"] = seed
    return p


def test_initial_states_golden(amd):
    params = amd.CoreParams(rng_problem(GOLDEN["seed"]))
    st = amd.CoreState(params, 8, stream_id=0)
    np.testing.assert_array_equal(st.rng_words().ravel(), GOLDEN["host_stream0_8slots_flattened"])
    st1 = amd.CoreState(params, 8, stream_id=1)
    np.testing.assert_array_equal(st1.rng_words().ravel()[:8], GOLDEN["host_stream1_first8"])


def test_reseed_golden(amd):
    g = GOLDEN["reseed"]
    params = amd.CoreParams(rng_problem(GOLDEN["seed"]))
    st = amd.CoreState(params, g["size"])
    st.reseed(g["event_id"])
    draws = st.rng_draw(1, "u32")[:, 0]
    assert list(draws[:: g["stride"]]) == g["expected_first_draw"]


@pytest.mark.parametrize("num_slots", [4096, 1 << 20])
def test_reseed_matches_oracle(amd, num_slots):
    """Full reseed at the bench state size, checked on a sample of slots."""
    params = amd.CoreParams(rng_problem(20220511))
    st = amd.CoreState(params, num_slots)
    event = 123457
    st.reseed(event)
    words = st.rng_words()
    rng = np.random.default_rng(1)
    sample = np.unique(np.concatenate([[0, num_slots - 1], rng.integers(0, num_slots, 512)]))
    expected = ob.rng_initialize_batch(20220511, event * num_slots + sample,
                                       np.zeros(len(sample), dtype=np.uint64))
    np.testing.assert_array_equal(words[sample], expected)


def test_random_subsequence_offsets(amd):
    n = 4096
    params = amd.CoreParams(rng_problem(77))
    st = amd.CoreState(params, n)
    rng = np.random.default_rng(5)
    sub = rng.integers(0, 2**63, n, dtype=np.uint64)
    off = rng.integers(0, 2**40, n, dtype=np.uint64)
    off[:16] = np.arange(16)
    st.rng_initialize(sub, off)
    expected = ob.rng_initialize_batch(77, sub, off)
    np.testing.assert_array_equal(st.rng_words(), expected)


@pytest.mark.parametrize("kind", ["u32", "f64", "f32"])
def test_draws_bit_exact(amd, kind):
    n = 1 << 16
    params = amd.CoreParams(rng_problem(12345))
    st = amd.CoreState(params, n)
    start = st.rng_words()
    gpu = st.rng_draw(32, kind)
    cpu, end = ob.rng_draw_batch(start, 32, kind)
    np.testing.assert_array_equal(gpu.view(np.uint32 if kind != "f64" else np.uint64),
                                  cpu.view(np.uint32 if kind != "f64" else np.uint64))
    np.testing.assert_array_equal(st.rng_words(), end)
