"""Structural facts the Barnes-Hut child blocks rely on, checked on the CPU.

The device builds a Karras binary radix tree over the sorted 63-bit Morton keys
(identical keys ordered by index), marks as octree cells ("kept") the nodes whose
common prefix reaches a deeper 3-bit group than their parent's, and has every cell
below the bucket level gather its children by descending at most three binary
levels (`block_kernel`, csrc/bh.hip).  Here the same tree is built in plain Python
and the descent's result is compared, cell by cell, with an independent definition
of a compressed octree's children: the cell's particles grouped by their next
3-bit group; a group of one particle is that particle, a larger group is the
deepest cell holding exactly that group (reference child order x + 2y + 4z,
src/Sim/Octree.cpp:25-41, is key order).
"""
import numpy as np
import pytest

LEVELS = 21
BUCKET = LEVELS


def clz64(x):
    return 64 - int(x).bit_length()


def clz32(x):
    return 32 - int(x).bit_length()


def delta(keys, i, j):
    n = len(keys)
    if j < 0 or j >= n:
        return -1
    if keys[i] == keys[j]:
        return 64 + clz32(i ^ j)
    return clz64(keys[i] ^ keys[j])


def level_of(d):
    return BUCKET if d >= 64 else (d - 1) // 3


def radix_tree(keys):
    """Karras (2012): for each internal node its range, split and children
    (a child < 0 is the particle -c - 2), as the device builds it."""
    n = len(keys)
    m = n - 1
    child = [None] * m
    first = [0] * m
    last = [0] * m
    level = [0] * m
    for i in range(m):
        d = 1 if delta(keys, i, i + 1) - delta(keys, i, i - 1) >= 0 else -1
        dmin = delta(keys, i, i - d)
        lmax = 2
        while delta(keys, i, i + lmax * d) > dmin:
            lmax <<= 1
        l = 0
        t = lmax >> 1
        while t >= 1:
            if delta(keys, i, i + (l + t) * d) > dmin:
                l += t
            t >>= 1
        j = i + l * d
        dnode = delta(keys, i, j)
        s = 0
        t = l
        while True:
            t = (t + 1) >> 1
            if delta(keys, i, i + (s + t) * d) > dnode:
                s += t
            if t <= 1:
                break
        gamma = i + s * d + min(d, 0)
        lo, hi = min(i, j), max(i, j)
        lc = -gamma - 2 if lo == gamma else gamma
        rc = -(gamma + 1) - 2 if hi == gamma + 1 else gamma + 1
        child[i] = (lc, rc)
        first[i], last[i] = lo, hi
        level[i] = level_of(dnode)
    parent = [-1] * m
    for i in range(m):
        for c in child[i]:
            if c >= 0:
                parent[c] = i
    kept = [i == 0 or level[i] > level[parent[i]] for i in range(m)]
    return child, first, last, level, kept


def descent(i, child, kept):
    """The device's gather: children of cell i in key order, at most three
    binary levels down; also the deepest level used."""
    out = []
    deepest = 0
    for c1 in child[i]:
        if c1 < 0 or kept[c1]:
            out.append(c1)
            deepest = max(deepest, 1)
            continue
        for c2 in child[c1]:
            if c2 < 0 or kept[c2]:
                out.append(c2)
                deepest = max(deepest, 2)
                continue
            for c3 in child[c2]:
                assert c3 < 0 or kept[c3], "a fourth binary level would be needed"
                out.append(c3)
                deepest = max(deepest, 3)
    return out, deepest


def expected_children(i, keys, first, last, level):
    """Independent definition: the cell's particles grouped by their next
    3-bit group of the key, in key order."""
    f, l, L = first[i], last[i], level[i]
    shift = 3 * (LEVELS - (L + 1))
    groups = []
    for q in range(f, l + 1):
        g = int(keys[q]) >> shift
        if groups and groups[-1][0] == g:
            groups[-1][2] = q
        else:
            groups.append([g, q, q])
    node_of_range = {(first[k], last[k]): k for k in range(len(first))}
    out = []
    for _, gf, gl in groups:
        out.append(-gf - 2 if gf == gl else node_of_range[(gf, gl)])
    return out


def clustered_keys(rng, n, dup_fraction):
    """Morton-like 63-bit keys with clusters at several scales and a share of
    exact duplicates (buckets)."""
    centres = rng.integers(0, 1 << 63, 8, dtype=np.uint64)
    keys = []
    for _ in range(n):
        c = int(centres[rng.integers(0, len(centres))])
        span = 1 << int(rng.integers(3, 63))
        keys.append((c ^ int(rng.integers(0, span))) & ((1 << 63) - 1))
    keys = np.array(keys, dtype=np.uint64)
    ndup = int(dup_fraction * n)
    if ndup:
        keys[rng.integers(0, n, ndup)] = keys[rng.integers(0, n, ndup)]
    return np.sort(keys, kind="stable")


@pytest.mark.parametrize("seed,n,dup", [(1, 400, 0.0), (2, 1500, 0.05), (3, 1200, 0.3), (4, 64, 0.5)])
def test_block_descent_gathers_the_octree_children(seed, n, dup):
    rng = np.random.default_rng(seed)
    keys = [int(k) for k in clustered_keys(rng, n, dup)]
    child, first, last, level, kept = radix_tree(keys)
    cells = [i for i in range(n - 1) if kept[i] and level[i] < BUCKET]
    assert cells
    deepest = 0
    for i in cells:
        got, d = descent(i, child, kept)
        deepest = max(deepest, d)
        assert 2 <= len(got) <= 8, (i, got)
        assert got == expected_children(i, keys, first, last, level), i
    assert deepest == 3  # the bound is reached, not just respected


def test_kept_flag_from_the_parent_matches_the_childs_own_level():
    """The device's parent writes each internal child's level from the child's
    range end points; that equals the level the child's own thread finds."""
    rng = np.random.default_rng(7)
    keys = [int(k) for k in clustered_keys(rng, 800, 0.2)]
    child, first, last, level, kept = radix_tree(keys)
    for i in range(len(child)):
        for c in child[i]:
            if c >= 0:
                assert level_of(delta(keys, first[c], last[c])) == level[c]
                assert kept[c] == (level[c] > level[i])
