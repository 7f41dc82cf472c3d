"""The re-sort's position formulas (vkfluidsim_amd/csrc/sort.hip, re-sort
section) against the stable sort they replace, on the CPU with numpy: for an
ascending previous key array `old`, its bucket starts `cs` (first index with
old >= k) and any new keys, the changed entries (key differs from old at the
same index) are ranked among themselves and every entry gets

    unchanged j:      (j - nb[j]) + lb[key_j] + #{changed b : key_b = key_j, b < j}
    changed, rank r:  (c - nb[c]) + r,  c = clamp(b, cs[key_b], cs[key_b + 1])

(nb = changed entries before an index, lb[k] = changed keys below k; the
changed entries of key k are ranks lb[k] .. lb[k + 1] - 1, in index order),
which must equal the position of the entry in np.argsort(keys,
kind="stable"); the updated bucket starts (cs[k] - nb[cs[k]] + lb[k]) must
equal the bucket starts of the new keys.  A dropped term, a swapped clamp bound or an unstable tie breaks
one of the cases (few, all, none changed; duplicates; a table of one key)."""
import numpy as np
import pytest


def _resort(old, keys, table):
    n = len(keys)
    cs = np.searchsorted(old, np.arange(table + 1), side="left")
    changed = keys != old
    nb = np.concatenate([[0], np.cumsum(changed)])
    b = np.flatnonzero(changed)
    order = np.lexsort((b, keys[b]))          # (key, index) order of the changed entries
    bks, bis = keys[b][order], b[order]
    out = np.empty(n, np.int64)
    r = np.arange(len(bis))
    c = np.clip(bis, cs[bks], cs[bks + 1])
    out[bis] = (c - nb[c]) + r
    lb = np.searchsorted(bks, np.arange(table + 1), side="left")
    u = np.flatnonzero(~changed)
    lo, hi = lb[keys[u]], lb[keys[u] + 1]
    same_key_before = np.array([np.searchsorted(bis[a:z], j) for a, z, j in zip(lo, hi, u)], np.int64)
    out[u] = (u - nb[u]) + lo + same_key_before
    new_cs = (cs - nb[cs]) + lb
    return out, new_cs


CASES = [(1, 0.5, 1), (7, 0.3, 7), (100, 0.05, 100), (1000, 0.01, 1000), (1000, 0.5, 50), (5000, 0.002, 5000),
         (3000, 1.0, 3000), (2000, 0.0, 2000), (500, 0.2, 1)]


@pytest.mark.parametrize("n,frac,table", CASES)
def test_resort_positions_equal_the_stable_sort(n, frac, table):
    rng = np.random.default_rng(n * 31 + table)
    for shuffled in (False, True):
        for _ in range(10):
            old = np.sort(rng.integers(0, table, n))
            keys = rng.integers(0, table, n) if shuffled else old.copy()
            ch = rng.random(n) < frac
            keys[ch] = rng.integers(0, table, int(ch.sum()))
            out, new_cs = _resort(old, keys, table)
            perm = np.empty(n, np.int64)
            perm[out] = np.arange(n)
            np.testing.assert_array_equal(perm, np.argsort(keys, kind="stable"))
            np.testing.assert_array_equal(new_cs, np.searchsorted(np.sort(keys), np.arange(table + 1), side="left"))


def test_a_wrong_clamp_is_caught():
    """Clamping to the new instead of the previous bucket starts misplaces changed entries."""
    rng = np.random.default_rng(5)
    old = np.sort(rng.integers(0, 50, 400))
    keys = old.copy()
    ch = rng.random(400) < 0.3
    keys[ch] = rng.integers(0, 50, int(ch.sum()))
    out, _ = _resort(old, keys, 50)
    # the same positions computed with the new keys' bucket starts
    cs_new = np.searchsorted(np.sort(keys), np.arange(51), side="left")
    changed = keys != old
    nb = np.concatenate([[0], np.cumsum(changed)])
    b = np.flatnonzero(changed)
    order = np.lexsort((b, keys[b]))
    bks, bis = keys[b][order], b[order]
    c = np.clip(bis, cs_new[bks], cs_new[bks + 1])
    wrong = (c - nb[c]) + np.arange(len(bis))
    assert not np.array_equal(wrong, out[bis])


def _rank_by_counting(bkeys, waves=16):
    """k_resort_rank's formula: the entry at compacted position p goes to
    #{q : (key_q, q) < (key_p, p)}, the comparisons split over `waves` equal
    shares of q whose partial counts are summed."""
    c = len(bkeys)
    comp = bkeys.astype(np.int64) * (1 << 32) + np.arange(c)
    chunk = (c + waves - 1) // waves
    rank = np.zeros(c, np.int64)
    for w in range(waves):
        q0, q1 = min(c, w * chunk), min(c, w * chunk + chunk)
        rank += (comp[None, q0:q1] < comp[:, None]).sum(axis=1)
    return rank


@pytest.mark.parametrize("c,table", [(1, 5), (16, 3), (100, 1000), (1000, 40), (4097, 300), (16384, 1 << 20)])
def test_rank_by_counting_is_the_stable_order(c, table):
    """Compacted changed keys (index order) placed by their counting rank form
    the (key, index) order the merge expects: np.lexsort's, ties by position."""
    rng = np.random.default_rng(c)
    bkeys = rng.integers(0, table, c).astype(np.uint32)
    rank = _rank_by_counting(bkeys)
    order = np.empty(c, np.int64)
    order[rank] = np.arange(c)
    np.testing.assert_array_equal(order, np.lexsort((np.arange(c), bkeys)))
    # ties broken the other way (later position first) would not be stable
    if len(np.unique(bkeys)) < c:
        comp = bkeys.astype(np.int64) * (1 << 32) + (c - 1 - np.arange(c))
        wrong = (comp[None, :] < comp[:, None]).sum(axis=1)
        assert not np.array_equal(wrong, rank)
