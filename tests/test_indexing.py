"""Index algebra: closed forms vs the reference's enumeration (src/Utils.jl:10-19,
src/Basis.jl:203) and its KATs (test/runtests.jl:17-24)."""
import numpy as np
import pytest

from mi_amd import indexing as I
from oracle import oracle as O


def test_reference_utils_kats():
    from mi_amd.basis import factorial2
    assert factorial2(6) == 48
    assert factorial2(-1) == 1
    assert I.ipairs(3) == [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3)]
    assert I.iindex(1, 1, 1, 1) == 1
    assert I.iindex(1, 1, 1, 2) == I.iindex(1, 1, 2, 1) == I.iindex(1, 2, 1, 1) == I.iindex(2, 1, 1, 1) == 2
    assert I.iindex(1, 1, 2, 2) == I.iindex(2, 2, 1, 1) == 4


@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 30])
def test_chrr_order_is_enumeration(n):
    quartets = I.iiterator(n)
    assert len(quartets) == I.n_unique(n) == O.n_unique(n)
    for pos, (i, j, k, l) in enumerate(quartets, start=1):
        assert I.chrr_index(n, i, j, k, l) == pos
    # vectorised forms agree
    i, j, k, l = I.canonical_quartets(n)
    assert np.array_equal(I.slots(n, i, j, k, l, 1), np.arange(I.n_unique(n)))


@pytest.mark.parametrize("n", [1, 2, 5, 16, 58])
def test_iindex_is_permutation(n):
    i, j, k, l = I.canonical_quartets(n)
    s = I.slots(n, i, j, k, l, 0)
    assert np.array_equal(np.sort(s), np.arange(I.n_unique(n)))
    for q in range(0, len(i), max(1, len(i) // 50)):
        assert s[q] + 1 == I.iindex(int(i[q]), int(j[q]), int(k[q]), int(l[q]))
        assert O.lib().oracle_slot(n, int(i[q]), int(j[q]), int(k[q]), int(l[q]), 0) == s[q]
        assert O.lib().oracle_slot(n, int(i[q]), int(j[q]), int(k[q]), int(l[q]), 1) == q


def test_orders_differ_as_documented():
    # D4: both drivers write the same slots in different permutations
    n = 16
    i, j, k, l = I.canonical_quartets(n)
    a, b = I.slots(n, i, j, k, l, 0), I.slots(n, i, j, k, l, 1)
    assert not np.array_equal(a, b)
    assert np.count_nonzero(a == b) < 10
