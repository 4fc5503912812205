"""celer_hip_partition_keys_noalloc (static scratch, no allocation) against
celer_hip_partition_keys (scratch allocated per call): identical results on
two calls in a row, the second smaller than the first (the static table
then holds the first call's larger histogram), plus the limit check."""

import ctypes

import numpy as np
import pytest

from tests.hip_buffers import DeviceBuffer, synchronize
from tests.test_gpu_partition import expected, partition

pytestmark = pytest.mark.gpu


def test_noalloc_matches_allocating_entry_point(amd):
    lib = amd.lib()
    rng = np.random.default_rng(77)
    for n, num_keys in (((1 << 20) + 3, 33), (5000, 9)):
        keys = rng.integers(0, 40, n)
        rc_a, sorted_a, off_a = partition(lib, "celer_hip_partition_keys", keys, num_keys)
        rc_b, sorted_b, off_b = partition(lib, "celer_hip_partition_keys_noalloc", keys,
                                          num_keys)
        assert rc_a == 0 and rc_b == 0
        np.testing.assert_array_equal(sorted_b, sorted_a)
        np.testing.assert_array_equal(off_b, off_a)
        order, offsets = expected(keys, num_keys)
        np.testing.assert_array_equal(sorted_b, order)
        np.testing.assert_array_equal(off_b, offsets)


def test_noalloc_rejects_sizes_beyond_the_limit(amd):
    lib = amd.lib()
    f = lib.celer_hip_partition_keys_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p]
    d_offsets = DeviceBuffer(4 * 3)
    # Nothing is read or written when the size is rejected
    rc = f(None, (1 << 22) + 1, 2, None, d_offsets.ptr, None)
    synchronize()
    assert rc != 0
    assert b"allocation-free limit" in lib.celer_amd_last_error()
