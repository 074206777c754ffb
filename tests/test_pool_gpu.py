"""Device-array pool: freed blocks are reused by size class and counted
as free memory; release_pool hands them back to the device."""

import numpy as np
import pytest

from dyfox_amd import _native as N

pytestmark = pytest.mark.gpu


def test_freed_block_is_reused_for_the_same_size_class(dev):
    a = N.DeviceArray.empty((1000,), np.float64)
    p = a.ptr
    del a
    b = N.DeviceArray.empty((999,), np.float64)        # same 8 KiB class
    assert b.ptr == p
    c = N.DeviceArray.empty((1000,), np.float64)       # class now empty
    assert c.ptr != p


def test_pooled_bytes_count_as_free_and_release(dev):
    N.release_pool()
    big = N.DeviceArray.empty((64 << 20,), np.uint8)
    f_held = N.mem_info()[0]
    del big
    assert N._pool_bytes == 64 << 20
    assert N.mem_info()[0] >= f_held + (64 << 20) - (2 << 20)
    N.release_pool()
    assert N._pool_bytes == 0 and not N._pool


def test_reused_block_carries_data_written_through_it(dev):
    x = np.arange(4096, dtype=np.float64)
    a = N.DeviceArray.from_host(x)
    del a
    b = N.DeviceArray.from_host(-x)
    assert np.array_equal(b.get(), -x)
