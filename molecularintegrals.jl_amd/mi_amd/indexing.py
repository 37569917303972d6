"""Unique-quartet index algebra of the reference (src/Utils.jl:10-19,
src/Basis.jl:201-212), as closed forms over 1-based function indices.

* ``iindex(i,j,k,l)`` = triangle(triangle(i,j), triangle(k,l)) -- the slot of
  ``all_twoe_ints`` (src/ERI.jl:109).
* ``chrr_index(n,i,j,k,l)`` -- the slot ``all_twoe_ints_chrr`` writes, i.e. the
  position of (i,j,k,l) in ``enumerate(iiterator(n))`` (src/Basis.jl:203):
  with M = n(n+1)/2 and ij <= kl,  (ij-1)(M+1) - ij(ij-1)/2 + kl - ij + 1.
"""
import numpy as np

ORDER_IINDEX = 0
ORDER_CHRR = 1


def triangle(i, j=None):
    if j is None:
        return i * (i + 1) // 2
    return triangle(j - 1) + i if i < j else triangle(i - 1) + j


def iindex(i, j, k, l):
    return triangle(triangle(i, j), triangle(k, l))


def n_unique(n):
    return n * (n + 1) * (n * n + n + 2) // 8


def chrr_index(n, i, j, k, l):
    ij, kl = triangle(i, j), triangle(k, l)
    if ij > kl:
        ij, kl = kl, ij
    M = n * (n + 1) // 2
    return (ij - 1) * (M + 1) - ij * (ij - 1) // 2 + kl - ij + 1


def ipairs(n):
    return [(i, j) for i in range(1, n + 1) for j in range(1, i + 1)]


def iiterator(n):
    """Canonical quartets in the reference's enumeration order (src/Utils.jl:17)."""
    P = ipairs(n)
    return [(i, j, k, l) for (i, j) in P for (k, l) in P if triangle(i, j) <= triangle(k, l)]


def canonical_quartets(n):
    """Vectorised (i,j,k,l) arrays (1-based) in enumeration order."""
    ii, jj = np.tril_indices(n)
    ii = ii + 1
    jj = jj + 1  # row-major over i then j == ipairs order
    M = len(ii)
    p, q = np.triu_indices(M)  # p <= q, row-major over p == iiterator order
    return ii[p], jj[p], ii[q], jj[q]


def slots(n, i, j, k, l, order):
    """0-based slots of canonical quartets (numpy int64 arrays)."""
    i, j, k, l = (np.asarray(x, dtype=np.int64) for x in (i, j, k, l))
    ij = np.where(i >= j, i * (i - 1) // 2 + j, j * (j - 1) // 2 + i)
    kl = np.where(k >= l, k * (k - 1) // 2 + l, l * (l - 1) // 2 + k)
    lo, hi = np.minimum(ij, kl), np.maximum(ij, kl)
    if order == ORDER_IINDEX:
        return hi * (hi - 1) // 2 + lo - 1
    M = n * (n + 1) // 2
    return (lo - 1) * (M + 1) - lo * (lo - 1) // 2 + hi - lo
