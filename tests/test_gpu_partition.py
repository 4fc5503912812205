"""celer_hip_partition_keys: the stable multi-split of sort_tracks over a
caller's keys, against its closed form (numpy's stable argsort of the
clamped keys and the cumulative key counts).

The reference sorts track slots by per-slot action keys with
thrust::sort_by_key (track/detail/TrackSortUtils.cu:145-206); the split
here is stable, so the permutation is fixed exactly. Cases: empty and
single-element inputs, sizes around the 1024-index tile, a million
indices, one key, 64 keys, keys beyond the last (clamped), constant,
ascending and descending keys.
"""

import ctypes

import numpy as np
import pytest

from tests.hip_buffers import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu


def expected(keys, num_keys):
    k = np.minimum(keys.astype(np.int64), num_keys - 1)
    order = np.argsort(k, kind="stable").astype(np.uint32)
    offsets = np.concatenate([[0], np.cumsum(np.bincount(k, minlength=num_keys))])
    return order, offsets.astype(np.uint32)


def partition(lib, name, keys, num_keys):
    f = getattr(lib, name)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p]
    n = len(keys)
    d_keys = DeviceBuffer.from_array(keys.astype(np.uint8))
    d_sorted = DeviceBuffer(4 * n)
    d_offsets = DeviceBuffer(4 * (num_keys + 1))
    rc = f(d_keys.ptr, n, num_keys, d_sorted.ptr, d_offsets.ptr, None)
    synchronize()
    return rc, d_sorted.to_array(np.uint32, n), d_offsets.to_array(np.uint32, num_keys + 1)


CASES = [
    ("empty", 0, 5, "random"),
    ("single", 1, 3, "random"),
    ("one key", 3000, 1, "random"),
    ("tile", 1024, 7, "random"),
    ("tile+1", 1025, 7, "random"),
    ("two tiles", 2048, 7, "random"),
    ("two tiles+1", 2049, 7, "random"),
    ("clamped", 10_000, 17, "wide"),
    ("constant", 5000, 64, "constant"),
    ("ascending", 4097, 64, "ascending"),
    ("descending", 4097, 64, "descending"),
    ("million", (1 << 20) + 7, 29, "random"),
]


def make_keys(n, num_keys, kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "random":
        return rng.integers(0, num_keys, n)
    if kind == "wide":
        return rng.integers(0, 256, n)
    if kind == "constant":
        return np.full(n, num_keys // 2)
    k = np.arange(n) * num_keys // max(n, 1)
    return k if kind == "ascending" else k[::-1].copy()


@pytest.mark.parametrize("label,n,num_keys,kind", CASES)
def test_partition_keys(amd, label, n, num_keys, kind):
    keys = make_keys(n, num_keys, kind, seed=n + num_keys)
    rc, got, off = partition(amd.lib(), "celer_hip_partition_keys", keys, num_keys)
    assert rc == 0
    order, offsets = expected(keys, num_keys)
    np.testing.assert_array_equal(off, offsets, err_msg=label)
    np.testing.assert_array_equal(got, order, err_msg=label)


def test_partition_keys_rejects_bad_key_counts(amd):
    f = amd.lib().celer_hip_partition_keys
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p]
    for nk in (0, 65):
        assert f(None, 0, nk, None, None, None) != 0
