"""The Barnes-Hut walk's stack and interaction list never overflow, checked on
the CPU.

`traverse_kernel` (csrc/bh.hip) pops a batch of up to eight cells from the top of
a wave's LDS stack of `kStackCap` entries and pushes, in their place, the opened
children (at most eight per cell, so each popped cell nets at most seven
entries).  The batch size is

    kmax = min(8, sp, max((kStackCap - kDfsReserve - sp) / 7, 1))

possibly lowered further (a bucket inside the batch, room in the interaction
list), never raised.  The argument that this cannot overflow:

* a batch of k >= 1 taken while (kStackCap - kDfsReserve - sp) / 7 >= 1 leaves
  sp + 7k <= H0 = kStackCap - kDfsReserve, so every entry at a position >= H0
  is written by a depth-first step (k = 1), whose children land at positions
  t .. t+7 above the popped entry at position t;
* children sit at a strictly deeper octree level than their cell (the kept
  cells of the radix tree), so by induction over the writes the entry at
  position i >= H0 has level >= floor((i - H0) / 7) + 1;
* only cells of level <= kLevels - 1 push anything (a level-kLevels bucket is
  expanded particle by particle), so the highest position popped with pushes
  is H0 + 7 (kLevels - 1) - 1 and the stack peaks at most at
  H0 + 7 kLevels <= kStackCap.

The test parses the constants and the batch rule from the kernel source, checks
the arithmetic, and runs the same pop/push discipline over adversarial trees
(full octrees that fill the stack in batches, topped by deep combs that then
descend depth-first), where the bound must hold and be approached; the same
trees overflow a reserve sized for fourteen levels, so the simulation can fail.

The interaction list (`kListCap` entries per wave) takes at most eight entries
per cell of a batch and one per particle of a bucket; its guards (flush or a
smaller batch) are read from the kernel source and checked for every list
length and batch size, together with the flush's padding to a multiple of four.
"""
import pathlib
import re

import numpy as np
import pytest

SRC = pathlib.Path(__file__).resolve().parents[1] / "procedural-universe_amd" / "csrc" / "bh.hip"


def _int_expr(text):
    """A product of integer literals, e.g. `7 * 22`."""
    parts = [p.strip() for p in text.split("*")]
    assert all(re.fullmatch(r"\d+", p) for p in parts), text
    out = 1
    for p in parts:
        out *= int(p)
    return out


def kernel_constants():
    src = SRC.read_text()
    cap = _int_expr(re.search(r"constexpr int kStackCap = ([^;]+);", src).group(1))
    reserve = _int_expr(re.search(r"constexpr int kDfsReserve = ([^;]+);", src).group(1))
    levels = _int_expr(re.search(r"constexpr int kLevels = ([^;]+);", src).group(1))
    rule = re.search(r"const int kmax = ([^;]+);", src).group(1)
    return cap, reserve, levels, rule


def list_guards():
    """kListCap and the two guards in front of list appends."""
    src = SRC.read_text()
    cap = _int_expr(re.search(r"constexpr int kListCap = ([^;]+);", src).group(1))
    bucket = re.search(r"if \(cnt > kListCap - (\d+)\) flush\(\);", src)
    batch = re.search(r"if \(cnt > kListCap - 8 \* k\) \{\s*const int fit = \(kListCap - cnt\) >> 3;\s*"
                      r"if \(fit >= (\d+)\) k = fit;\s*else flush\(\);", src)
    assert bucket and batch, "list guards not found in the kernel source"
    return cap, int(bucket.group(1)), int(batch.group(1))


def proof_bound(cap, reserve, levels):
    """Largest stack height the argument above allows."""
    h0 = cap - reserve
    return max(h0, h0 + 7 * levels)


BUCKET = "bucket"


def children(node, levels):
    """Lazy adversarial trees.  A node is (level, kind, arg):
    ("full", d): eight children, full octree for d more levels, then combs;
    ("comb", (c, s)): eight children, the one at slot s continues for c more
    levels, the others are cells whose children are particles only;
    ("leafcell", _): no cell children.  The deepest comb cell's continuing
    child is a bucket at level `levels`."""
    level, kind, arg = node
    if kind == BUCKET or kind == "leafcell":
        return []
    if kind == "full":
        d, comb, slot = arg
        if d > 1:
            return [(level + 1, "full", (d - 1, comb, slot))] * 8
        return [(level + 1, "comb", (comb, slot))] * 8
    c, slot = arg
    out = [(level + 1, "leafcell", None)] * 8
    if c > 1:
        out[slot] = (level + 1, "comb", (c - 1, slot))
    else:
        out[slot] = (levels, BUCKET, None)
    return out


def walk(root, cap, reserve, levels, rng=None):
    """The kernel's stack discipline with every child opened.  Returns the
    largest height reached and whether a push went past `cap`."""
    stack = [root]
    peak = 1
    while stack:
        sp = len(stack)
        kmax = min(8, sp, max((cap - reserve - sp) // 7, 1))
        top = stack[-kmax:][::-1]  # j = 0 is the top entry
        k = kmax
        for j, e in enumerate(top):
            if e[1] == BUCKET:
                k = j
                break
        if k == 0:
            stack.pop()  # a bucket alone: its particles go to the list
            continue
        if rng is not None and k > 4 and rng.random() < 0.3:
            k = int(rng.integers(4, k + 1))  # list room lowers the batch
        batch = top[:k]
        del stack[sp - k:]
        pushed = []
        for cell in batch:
            kids = children(cell, levels)
            for kid in kids:
                assert kid[0] > cell[0], "a child must sit deeper than its cell"
                assert kid[0] <= levels
            pushed.extend(kids)
        # the top cell's first child ends on top
        stack.extend(reversed(pushed))
        if len(stack) > cap:
            return max(peak, len(stack)), True
        peak = max(peak, len(stack))
    return peak, False


def adversarial_roots(levels):
    """Full octrees of depth d (filling the stack through batches) whose
    bottom cells start combs down to the bucket level: the combs then descend
    depth-first with seven siblings left behind per level."""
    for d in (2, 3, 4):
        for slot in (0, 3, 7):
            yield (0, "full", (d, levels - d, slot))


def test_batch_rule_in_the_kernel_is_the_one_checked():
    _, _, _, rule = kernel_constants()
    assert re.sub(r"\s+", "", rule) == "min(min(8,sp),max((kStackCap-kDfsReserve-sp)/7,1))"


def test_reserve_covers_a_depth_first_descent_through_every_level():
    cap, reserve, levels, _ = kernel_constants()
    bound = proof_bound(cap, reserve, levels)
    assert bound <= cap, (bound, cap)
    assert reserve >= 7 * levels


@pytest.mark.parametrize("root", list(adversarial_roots(21)), ids=lambda r: f"full{r[2][0]}_slot{r[2][2]}")
def test_walk_stays_under_the_bound(root):
    cap, reserve, levels, _ = kernel_constants()
    assert levels == 21
    bound = proof_bound(cap, reserve, levels)
    peak, over = walk(root, cap, reserve, levels)
    assert not over
    assert peak <= bound
    if root[2][2] == 0:
        # the continuing child on top: the descent really runs depth-first
        # above the batch region
        assert peak > cap - reserve, peak


def test_walk_with_lowered_batches_stays_under_the_bound():
    cap, reserve, levels, _ = kernel_constants()
    bound = proof_bound(cap, reserve, levels)
    rng = np.random.default_rng(11)
    for root in adversarial_roots(levels):
        peak, over = walk(root, cap, reserve, levels, rng)
        assert not over and peak <= bound


def test_the_same_trees_overflow_a_reserve_for_fewer_levels():
    """A reserve that covers a depth-first descent through fourteen levels
    only: batches then fill the stack too high for the combs below them."""
    cap, _, levels, _ = kernel_constants()
    results = [walk(root, cap, 7 * 14, levels) for root in adversarial_roots(levels)]
    assert any(over for _, over in results)


def padded(cnt):
    return cnt + ((4 - (cnt & 3)) & 3)


def test_list_guards_keep_every_append_inside_the_list():
    cap, bucket_margin, min_fit = list_guards()
    assert cap % 4 == 0
    reachable = set()
    for cnt in range(cap + 1):
        # a bucket's particles: one entry each, after the guard
        c = 0 if cnt > cap - bucket_margin else cnt
        assert c + 1 <= cap
        reachable.add(c + 1)
        # a batch of k cells: at most eight entries each, after the guard
        for k in range(1, 9):
            c, kk = cnt, k
            if c > cap - 8 * kk:
                fit = (cap - c) >> 3
                if fit >= min_fit:
                    kk = fit
                else:
                    c = 0
            assert 1 <= kk <= k
            assert c + 8 * kk <= cap, (cnt, k)
            reachable.update(range(c, c + 8 * kk + 1))
    # the flush pads with zero entries up to a multiple of four
    assert all(padded(c) <= cap for c in reachable)


def test_list_guards_without_the_bucket_margin_would_overflow():
    """The same exhaustive check catches a bucket guard that flushes only when
    the list is already full."""
    cap, _, _ = list_guards()
    assert any((0 if cnt > cap else cnt) + 1 > cap for cnt in range(cap + 1))
