"""GPU parity of the spatial-hash path: hash keys, stable sort, scan, offsets,
reorder -- bit-exact against the CPU oracle on identical inputs."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V
from vkfluidsim_amd import configs as CF

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 255, 256, 4095, 4096, 4097, 65537, 300_001])
def test_sort_pairs_random_keys(n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, n, n, dtype=np.uint32)
    sk, perm = V.sort_pairs(keys, max_value=n - 1)
    rk, rp = H.oracle_sort(keys, n - 1)
    np.testing.assert_array_equal(sk, rk)
    np.testing.assert_array_equal(perm, rp)


@pytest.mark.parametrize("case", ["same", "two", "sorted", "reversed", "nearly", "full32"])
def test_sort_pairs_structured(case):
    n = 200_003
    rng = np.random.default_rng(7)
    if case == "same":
        keys = np.full(n, 12345, np.uint32)
        maxv = n - 1
    elif case == "two":
        keys = rng.integers(0, 2, n).astype(np.uint32) * np.uint32(n - 1)
        maxv = n - 1
    elif case == "sorted":
        keys = np.sort(rng.integers(0, n, n).astype(np.uint32))
        maxv = n - 1
    elif case == "reversed":
        keys = np.sort(rng.integers(0, n, n).astype(np.uint32))[::-1].copy()
        maxv = n - 1
    elif case == "nearly":
        keys = np.sort(rng.integers(0, n, n).astype(np.uint32))
        idx = rng.integers(0, n, n // 50)
        keys[idx] = rng.integers(0, n, idx.size).astype(np.uint32)
        maxv = n - 1
    else:
        keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        maxv = 2**32 - 1
    sk, perm = V.sort_pairs(keys, max_value=maxv)
    ref_perm = np.argsort(keys, kind="stable").astype(np.uint32)
    np.testing.assert_array_equal(perm, ref_perm)
    np.testing.assert_array_equal(sk, keys[ref_perm])


@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 1_000_003])
def test_scan_exclusive(n):
    rng = np.random.default_rng(n + 11)
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) if n else np.zeros(0, np.uint32)
    out = V.scan_exclusive(vals)
    ref = (np.concatenate([[0], np.cumsum(vals.astype(np.uint64))[:-1]]) % 2**32).astype(np.uint32) if n else vals
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("n", [1, 5, 4099, 100_000])
def test_spatial_offsets(n):
    rng = np.random.default_rng(n)
    sk = np.sort(rng.integers(0, n, n).astype(np.uint32))
    np.testing.assert_array_equal(V.spatial_offsets(sk, init=True), H.oracle_offsets(sk))
    # init = false only writes the starts of present keys
    pre = np.full(n, 0xDEADBEEF, np.uint32)
    got = V.spatial_offsets(sk, init=False, prefill=pre)
    ref = H.oracle_offsets(sk)
    present = ref != n
    np.testing.assert_array_equal(got[present], ref[present])
    assert np.all(got[~present] == 0xDEADBEEF)


def test_reorder_f32x3():
    n = 77_777
    rng = np.random.default_rng(3)
    perm = rng.permutation(n).astype(np.uint32)
    src = rng.standard_normal((n, 3)).astype(np.float32)
    np.testing.assert_array_equal(V.reorder_f32x3(perm, src), src[perm])


def _hash_run(x, cell, stride):
    n = x.shape[0]
    if stride == 4:
        x4 = np.zeros((n, 4), np.float32)
        x4[:, :3] = x
        d = V.DeviceArray.from_host(x4)
    else:
        d = V.DeviceArray.from_host(np.ascontiguousarray(x, np.float32))
    h = V.SpatialHash(n, cell)
    h.run(d.ptr, stride)
    out = h.results()
    h.close()
    d.free()
    return out


@pytest.mark.parametrize("stride", [3, 4])
def test_spatial_hash_dam_break(stride):
    w = CF.dam_break()
    keys, perm, offs = _hash_run(w.x, w.h, stride)
    raw = H.oracle_keys(w.x, w.h, w.n)
    rk, rp = H.oracle_sort(raw, w.n - 1)
    np.testing.assert_array_equal(keys, rk)
    np.testing.assert_array_equal(perm, rp)
    np.testing.assert_array_equal(offs, H.oracle_offsets(rk))


def test_spatial_hash_negative_and_wrapping_cells():
    rng = np.random.default_rng(5)
    n = 50_000
    # cells from -60 to +80 per axis with cell size 0.1: covers the uint wrap below -25
    x = rng.uniform(-6.0, 8.0, (n, 3)).astype(np.float32)
    keys, perm, offs = _hash_run(x, 0.1, 3)
    raw = H.oracle_keys(x, 0.1, n)
    rk, rp = H.oracle_sort(raw, n - 1)
    np.testing.assert_array_equal(keys, rk)
    np.testing.assert_array_equal(perm, rp)
    np.testing.assert_array_equal(offs, H.oracle_offsets(rk))


def test_spatial_hash_uniform_8m():
    """Sort/neighbour stress workload: 8,388,608 particles uniform in 4 x 2 x 2, h = 0.02."""
    w = CF.uniform_stress()
    keys, perm, offs = _hash_run(w.x, w.h, 3)
    raw = H.oracle_keys(w.x, w.h, w.n)
    rk, rp = H.oracle_sort(raw, w.n - 1)
    assert np.array_equal(keys, rk)
    assert np.array_equal(perm, rp)
    assert np.array_equal(offs, H.oracle_offsets(rk))


def test_sort_pairs_beyond_2_to_the_28_keys():
    """The largest sizes: 2^28 + 12,345 keys (tile counts, indices and byte
    offsets past the particle limit of a simulation) with 16-bit keys, against
    numpy's stable sort of the same keys."""
    n = (1 << 28) + 12345
    print(f"case sort: {n} keys", flush=True)
    rng = np.random.default_rng(28)
    keys16 = rng.integers(0, 1 << 16, n, dtype=np.uint16)
    sk, perm = V.sort_pairs(keys16.astype(np.uint32), max_value=(1 << 16) - 1)
    ref_perm = np.argsort(keys16, kind="stable")
    assert np.array_equal(perm, ref_perm)
    assert np.array_equal(sk, keys16[ref_perm])
