"""vfs_sort_pairs_noalloc (fixed device scratch, no allocation) against
vfs_sort_pairs (caller workspace) through the C ABI: identical sorted keys and
permutations on consecutive calls, the second smaller than the first, and the
stable order of numpy's argsort."""
import ctypes as C

import numpy as np
import pytest

import vkfluidsim_amd as V
from vkfluidsim_amd._lib import DeviceArray, call

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20


def _sort_with_workspace(keys, max_value):
    n = keys.size
    ws = C.c_size_t(0)
    call("vfs_sort_workspace_size", n, max_value, C.byref(ws))
    dk = DeviceArray.from_host(keys)
    di = DeviceArray(4 * n)
    dw = DeviceArray(ws.value)
    call("vfs_sort_pairs", di.ptr, dk.ptr, n, max_value, dw.ptr, ws.value, None)
    call("vfs_stream_synchronize", None)
    return dk.to_host(np.uint32, n), di.to_host(np.uint32, n)


def _sort_noalloc(keys, max_value):
    n = keys.size
    dk = DeviceArray.from_host(keys)
    di = DeviceArray(4 * n)
    print(f"vfs_sort_pairs_noalloc n={n} max_value={max_value}", flush=True)
    call("vfs_sort_pairs_noalloc", di.ptr, dk.ptr, n, max_value, None)
    call("vfs_stream_synchronize", None)
    return dk.to_host(np.uint32, n), di.to_host(np.uint32, n)


def _check(keys, max_value):
    sk, perm = _sort_noalloc(keys, max_value)
    rk, rp = _sort_with_workspace(keys, max_value)
    np.testing.assert_array_equal(sk, rk)
    np.testing.assert_array_equal(perm, rp)
    ref = np.argsort(keys, kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(perm, ref)
    np.testing.assert_array_equal(sk, keys[ref])


def test_noalloc_sort_matches_workspace_sort_on_consecutive_calls():
    rng = np.random.default_rng(2024)
    # first call: the largest size, hash-table keys (3 passes of 8 bits)
    k1 = rng.integers(0, MAX_KEYS, MAX_KEYS, dtype=np.uint32)
    _check(k1, MAX_KEYS - 1)
    # second call, smaller and with full 32-bit keys (4 passes): the scratch
    # still holds the first call's words
    k2 = rng.integers(0, 2**32, 70_001, dtype=np.uint64).astype(np.uint32)
    _check(k2, 2**32 - 1)
    # third, tiny with many equal keys (one pass): stability
    k3 = rng.integers(0, 3, 257).astype(np.uint32)
    _check(k3, 2)


def test_noalloc_sort_rejects_more_keys_than_the_scratch_holds():
    n = MAX_KEYS + 1
    rc = V.lib.vfs_sort_pairs_noalloc(None, None, n, 0, None)
    assert rc == -1
    assert "exceeds" in V.lib.vfs_last_error().decode()
    assert V.lib.vfs_sort_pairs_noalloc(None, None, 0, 0, None) == 0
