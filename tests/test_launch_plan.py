"""The launch order of a grouped flush (bbjd_rec_plan, the planner that
bbjd_rec_flush runs on recorded launches; host logic only, no GPU): every
recorded launch of every solve goes out exactly once and in its solve's
order, a grouped launch holds launches of one kernel only (at most 16), a
closure (key 0) is never merged, and solves whose launch sequences differ in
the middle merge again in their common tail."""
import ctypes as C

import numpy as np

from bbjd_amd import _lib


def plan(lists, cap=100000):
    lib = _lib.load()
    lib.bbjd_rec_plan.restype = C.c_int
    lib.bbjd_rec_plan.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_int), C.c_int]
    lens = (C.c_int * max(1, len(lists)))(*[len(l) for l in lists])
    flat = [k for l in lists for k in l]
    keys = (C.c_int64 * max(1, len(flat)))(*flat)
    out = (C.c_int * cap)()
    nl = lib.bbjd_rec_plan(len(lists), lens, keys, out, cap)
    assert nl >= 0
    groups, i = [], 0
    for _ in range(nl):
        n = out[i]
        groups.append(list(out[i + 1:i + 1 + n]))
        i += 1 + n
    return groups


def check(lists, groups):
    """every launch once, in order, one key per group, closures alone;
    returns the number of closure launches"""
    pos = [0] * len(lists)
    closures = 0
    for g in groups:
        assert 1 <= len(g) <= 16 and len(set(g)) == len(g)
        ks = {lists[j][pos[j]] for j in g}
        assert len(ks) == 1, ks
        if 0 in ks:
            assert len(g) == 1
            closures += 1
        for j in g:
            pos[j] += 1
    assert pos == [len(l) for l in lists]
    return closures


def test_common_prefix_and_tail_merge():
    prefix, tail = [1, 2, 3, 4, 5], [20, 21, 22, 23, 24, 25]
    mids = [[10, 11, 12], [13, 14], [10, 15, 16, 17], [13, 14]]
    lists = [prefix + m + tail for m in mids]
    groups = plan(lists)
    check(lists, groups)
    # prefix and tail once each; the middles: 10 may merge for lists 0 and 2
    # only if they line up, 13 14 for lists 1 and 3
    assert len(groups) <= len(prefix) + len(tail) + 3 + 2 + 4
    assert sum(len(g) == 4 for g in groups) == len(prefix) + len(tail)


def test_closures_are_never_merged_and_keep_their_place():
    lists = [[1, 0, 2, 3], [1, 0, 2, 3], [0, 1, 2]]
    groups = plan(lists)
    assert check(lists, groups) == 3
    # the kernels after the closures still merge: 1 (three lists), 2 and 3
    assert len(groups) == 3 + 3


def test_random_lists_every_launch_once_in_order():
    rng = np.random.default_rng(11)
    for _ in range(200):
        nj = int(rng.integers(1, 20))
        lists = []
        for _ in range(nj):
            L = int(rng.integers(0, 30))
            lists.append([int(k) for k in rng.choice([0, 1, 2, 3, 4, 5, 6], size=L, p=[.1, .3, .2, .1, .1, .1, .1])])
        groups = plan(lists)
        check(lists, groups)
        assert len(groups) <= sum(len(l) for l in lists)
        assert len(groups) >= max([len(l) for l in lists] + [0])


def test_more_than_sixteen_identical_lists_split():
    lists = [[7, 8, 9]] * 20
    groups = plan(lists)
    check(lists, groups)
    assert len(groups) == 6 and sorted(len(g) for g in groups) == [4, 4, 4, 16, 16, 16]
