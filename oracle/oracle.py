"""ctypes front end of the CPU oracle (oracle/libhgp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by bench.py's CPU-baseline leg, always as the checker or the timed CPU baseline.
The product path (molecularintegrals.jl_amd) never imports this module.

Each wrapper names the reference function it restates (rpmuller/
MolecularIntegrals.jl file:line); the C source carries the details.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhgp_oracle.so")

LIT_D1, LIT_D2, LIT_D3, LIT_D5 = 1, 2, 4, 8
ORDER_IINDEX, ORDER_CHRR = 0, 1

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_Fm.restype = ctypes.c_double
        L.oracle_Fm.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_farray.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.oracle_farray_accurate.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.oracle_nao.argtypes = [ctypes.c_int]
        L.oracle_n_unique.restype = ctypes.c_int64
        L.oracle_n_unique.argtypes = [ctypes.c_int64]
        L.oracle_slot.restype = ctypes.c_int64
        L.oracle_slot.argtypes = [ctypes.c_int64] * 5 + [ctypes.c_int]
        vargs = [_dp, ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [_dp] * 4 + [ctypes.c_int]
        L.oracle_vrr.argtypes = vargs
        L.oracle_hrr.argtypes = [_dp] + [ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [_dp] * 4 + [ctypes.c_int]
        L.oracle_overlap_prim.restype = ctypes.c_double
        L.oracle_overlap_prim.argtypes = ([ctypes.c_double, _dp] + [ctypes.c_int] * 3) * 2
        L.oracle_pgbf_norm.restype = ctypes.c_double
        L.oracle_pgbf_norm.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_cgbf_norm.restype = ctypes.c_double
        L.oracle_cgbf_norm.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_coulomb_prim.restype = ctypes.c_double
        L.oracle_coulomb_prim.argtypes = ([ctypes.c_double, _dp] + [ctypes.c_int] * 3) * 4
        basis_args = [ctypes.c_int, _dp, _ip, _ip, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.oracle_eri_unique_hgp.argtypes = basis_args + [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_eri_unique_tho.argtypes = basis_args + [ctypes.c_int, _dp]
        L.oracle_eri_unique_hgp_ext.argtypes = basis_args + [ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_eri_unique_hgp_quad.argtypes = basis_args + [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.oracle_eri_slots_quad.argtypes = basis_args + [ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                                          ctypes.POINTER(ctypes.c_int64), ctypes.c_int, _dp]
        L.oracle_Fm_quad.restype = ctypes.c_double
        L.oracle_Fm_quad.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.oracle_eri_hgp_quartet_sample.restype = ctypes.c_int64
        L.oracle_eri_hgp_quartet_sample.argtypes = basis_args + [ctypes.c_int64, _ip, ctypes.c_int, _dp]
        L.oracle_jk_from_unique.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oracle_schwarz_pairs.argtypes = basis_args + [_dp]
        L.oracle_rys_roots.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.oracle_coulomb_rys_prim.restype = ctypes.c_double
        L.oracle_coulomb_rys_prim.argtypes = ([ctypes.c_double, _dp] + [ctypes.c_int] * 3) * 4
        L.oracle_eri_unique_rys.argtypes = basis_args + [ctypes.c_int, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def Fm(m, T):
    """Boys.jl:29-34."""
    return lib().oracle_Fm(int(m), float(T))


def farray(mmax, T):
    """Boys.jl:19-27."""
    out = np.zeros(mmax)
    lib().oracle_farray(int(mmax), float(T), out.ctypes.data_as(_dp))
    return out


def farray_accurate(mmax, T):
    """farray (Boys.jl:19-27) without its T >= 30 asymptotic branch: the Boys
    function for every T (what the corrected oracle and the GPU use)."""
    out = np.zeros(mmax)
    lib().oracle_farray_accurate(int(mmax), float(T), out.ctypes.data_as(_dp))
    return out


def nao(l):
    return lib().oracle_nao(int(l))


def _xyz(v):
    return _d(np.asarray(v, dtype=np.float64).reshape(3))


def vrr(amax, cmax, aexpn, bexpn, cexpn, dexpn, A, B, C, D, literal=0):
    """HGP.jl:371-376; returns vrrs[m, a, c] (0-based)."""
    mmax = amax + cmax + 1
    out = np.zeros((mmax, nao(amax), nao(cmax)))
    keep = [_xyz(x) for x in (A, B, C, D)]
    lib().oracle_vrr(out.ctypes.data_as(_dp), amax, cmax, aexpn, bexpn, cexpn, dexpn,
                     *[k[1] for k in keep], int(literal))
    return out


def hrr(la, lb, lc, ld, aexpn, bexpn, cexpn, dexpn, A, B, C, D, literal=0):
    """HGP.jl:392-424; returns hrrs[a, b, c, d] (0-based cumulative AO axes)."""
    out = np.zeros((nao(la + lb), nao(lb), nao(lc + ld), nao(ld)))
    keep = [_xyz(x) for x in (A, B, C, D)]
    lib().oracle_hrr(out.ctypes.data_as(_dp), la, lb, lc, ld, aexpn, bexpn, cexpn, dexpn,
                     *[k[1] for k in keep], int(literal))
    return out


def pgbf_norm(expn, I, J, K):
    return lib().oracle_pgbf_norm(float(expn), I, J, K)


def cgbf_norm(expns, coefs, I, J, K):
    e, ep = _d(expns)
    c, cp = _d(coefs)
    return lib().oracle_cgbf_norm(len(e), ep, cp, I, J, K)


def coulomb_prim(a, A, aIJK, b, B, bIJK, c, C, cIJK, d, D, dIJK):
    """ERI.jl:15-46 (unnormalised primitive THO)."""
    keep = [_xyz(x) for x in (A, B, C, D)]
    args = []
    for e, k, ijk in zip((a, b, c, d), keep, (aIJK, bIJK, cIJK, dIJK)):
        args += [float(e), k[1], *[int(x) for x in ijk]]
    return lib().oracle_coulomb_prim(*args)


def _basis_args(fb):
    arrs = fb.arrays()
    keep = [_d(arrs[0]), _i(arrs[1]), _i(arrs[2]), _d(arrs[3]), _d(arrs[4]), _d(arrs[5]), _d(arrs[6])]
    args = [fb.nshell, keep[0][1], keep[1][1], keep[2][1], keep[3][1], keep[4][1], keep[5][1], fb.nbf, keep[6][1]]
    return keep, args


def n_unique(nbf):
    return lib().oracle_n_unique(int(nbf))


def eri_unique_hgp(fb, order=ORDER_CHRR, literal=0, nthreads=0):
    """all_twoe_ints_chrr (HGP.jl:22-34) with D1-D3 and D5 (the T >= 30 Boys
    branch, Boys.jl:21-26) corrected unless literal."""
    keep, args = _basis_args(fb)
    out = np.zeros(n_unique(fb.nbf))
    rc = lib().oracle_eri_unique_hgp(*args, int(order), int(literal), int(nthreads), out.ctypes.data_as(_dp))
    if rc:
        raise ValueError(f"oracle_eri_unique_hgp rejected the basis (code {rc})")
    return out


def eri_unique_hgp_ext(fb, order=ORDER_CHRR, nthreads=0):
    """The same corrected HGP algorithm in x87 80-bit arithmetic: arbiter where
    the FP64 oracle loses digits to HRR cancellation."""
    keep, args = _basis_args(fb)
    out = np.zeros(n_unique(fb.nbf))
    rc = lib().oracle_eri_unique_hgp_ext(*args, int(order), int(nthreads), out.ctypes.data_as(_dp))
    if rc:
        raise ValueError(f"oracle_eri_unique_hgp_ext rejected the basis (code {rc})")
    return out


def eri_unique_hgp_quad(fb, order=ORDER_CHRR, accurate=False, nthreads=0):
    """The same HGP algorithm in binary128 (HGP.jl:50-242 with its recurrence
    rounding ~1e-34): the exact value of the reference algorithm.  accurate:
    Boys function evaluated exactly for every T (coulomb's Fgamma, ERI.jl:41)
    instead of farray's T >= 30 asymptotic branch (Boys.jl:21-26)."""
    keep, args = _basis_args(fb)
    out = np.zeros(n_unique(fb.nbf))
    rc = lib().oracle_eri_unique_hgp_quad(*args, int(order), int(bool(accurate)), int(nthreads),
                                          out.ctypes.data_as(_dp))
    if rc:
        raise ValueError(f"oracle_eri_unique_hgp_quad rejected the basis (code {rc})")
    return out


def eri_slots_quad(fb, slots, order=ORDER_CHRR, accurate=False, nthreads=0):
    """binary128 values (as eri_unique_hgp_quad) of the given 0-based slots only."""
    keep, args = _basis_args(fb)
    sl = np.ascontiguousarray(slots, dtype=np.int64)
    out = np.zeros(sl.size)
    rc = lib().oracle_eri_slots_quad(*args, int(order), int(bool(accurate)), sl.size,
                                     sl.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(nthreads),
                                     out.ctypes.data_as(_dp))
    if rc:
        raise ValueError(f"oracle_eri_slots_quad rejected the input (code {rc})")
    return out


def Fm_quad(m, T, accurate=True):
    """F_m(T) from the binary128 farray (Boys.jl:19-27); accurate: no T >= 30
    branch (the Boys function itself, Boys.jl:29-34)."""
    return lib().oracle_Fm_quad(int(m), float(T), int(bool(accurate)))


def eri_unique_tho(fb, order=ORDER_IINDEX):
    """all_twoe_ints(bfs, coulomb) (ERI.jl:104-112) -- tiny inputs only."""
    keep, args = _basis_args(fb)
    out = np.zeros(n_unique(fb.nbf))
    rc = lib().oracle_eri_unique_tho(*args, int(order), out.ctypes.data_as(_dp))
    if rc:
        raise ValueError(f"oracle_eri_unique_tho rejected the basis (code {rc})")
    return out


def rys_roots(n, X):
    """Roots(n, X) (Rys.jl:161-174): the n-point Rys rule from its definition
    (binary128 Golub-Welsch on the Boys moments); roots as u = t^2/(1-t^2)."""
    r = np.zeros(n)
    w = np.zeros(n)
    if lib().oracle_rys_roots(int(n), float(X), r.ctypes.data_as(_dp), w.ctypes.data_as(_dp)):
        raise ValueError(f"oracle_rys_roots(n={n}, X={X})")
    return r, w


def coulomb_rys_prim(a, A, aIJK, b, B, bIJK, c, C, cIJK, d, D, dIJK):
    """coulomb_rys on primitives (Rys.jl:59-84), without the primitive norms."""
    keep = [_xyz(x) for x in (A, B, C, D)]
    args = []
    for e, k, ijk in zip((a, b, c, d), keep, (aIJK, bIJK, cIJK, dIJK)):
        args += [float(e), k[1], *[int(x) for x in ijk]]
    return lib().oracle_coulomb_rys_prim(*args)


def eri_unique_rys(fb, order=ORDER_IINDEX, nthreads=0):
    """all_twoe_ints_rys (Rys.jl:16-25): every unique quartet by Rys quadrature."""
    keep, args = _basis_args(fb)
    out = np.zeros(n_unique(fb.nbf))
    rc = lib().oracle_eri_unique_rys(*args, int(order), int(nthreads), out.ctypes.data_as(_dp))
    if rc:
        raise ValueError(f"oracle_eri_unique_rys rejected the basis (code {rc})")
    return out


def eri_quartet_sample(fb, quartets, nthreads=0):
    keep, args = _basis_args(fb)
    q, qp = _i(np.asarray(quartets, dtype=np.int32).reshape(-1))
    cs = ctypes.c_double(0.0)
    n = lib().oracle_eri_hgp_quartet_sample(*args, len(q) // 4, qp, int(nthreads), ctypes.byref(cs))
    return n, cs.value


def jk_from_unique(nbf, eri_iindex, D):
    """J/K index algebra of make2JmK (ERI.jl:114-128)."""
    e, ep = _d(eri_iindex)
    d, dp = _d(D)
    J = np.zeros((nbf, nbf))
    K = np.zeros((nbf, nbf))
    lib().oracle_jk_from_unique(int(nbf), ep, dp, J.ctypes.data_as(_dp), K.ctypes.data_as(_dp))
    return J, K


def schwarz_pairs(fb):
    keep, args = _basis_args(fb)
    n = fb.nshell
    out = np.zeros(n * (n + 1) // 2)
    rc = lib().oracle_schwarz_pairs(*args, out.ctypes.data_as(_dp))
    if rc:
        raise ValueError("bad basis")
    return out
