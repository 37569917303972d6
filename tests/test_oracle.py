"""Pins the CPU oracle (oracle/hgp_oracle.c) before it is trusted as the checker.

What pins each oracle function (a plausible mistake must fail one of these):

* AO tables (Basis.jl:264-385)   -- test/runtests.jl:32-37 KATs (nao, m2ao) and
  the literal shell_indices rows (incl. f: xyy, xzz, xyz).
* Boys Fm / farray (Boys.jl)     -- KATs test/runtests.jl:70-72; mpmath closed
  form F_m(T) = 1F1(m+1/2; m+3/2; -T)/(2m+1) at 1e-14 relative over T in
  [0, 60]; F_0(T) = sqrt(pi/T) erf(sqrt T)/2; F_m(0) = 1/(2m+1); the
  downward-recursion identity; the T>=30 branch equals the reference's
  asymptotic formula exactly (e^-T dropped, as Boys.jl:21-26).  The last ulp of
  SpecialFunctions.gamma_inc is NOT reachable here: parity unpinned at that level.
* vrr (HGP.jl:135-242)           -- all 19 VRR KATs (test/runtests.jl:159-187).
* hrr (HGP.jl:392-424)           -- all 19 HRR KATs (test/runtests.jl:198-234).
* prefactor D1 / normalisation D2 / gather D3 -- closed-form s-primitive value
  2 pi^{5/2}/(zeta eta sqrt(zeta+eta)) exp(..) F_0 at unequal exponents (D1
  fails it by the factor zeta eta/(zeta+eta)); H2/STO-3G contracted KATs
  (runtests.jl:48-52; D2 fails them); independent THO restatement (ERI.jl:15-94)
  agreeing to 1e-13 on random s/p/d bases with general exponents (D3 and any
  dropped VRR/HRR term fail it).
* iindex / chrr slots            -- runtests.jl:22-24 KATs and brute-force
  enumeration (test_indexing.py).
* make2JmK index algebra (ERI.jl:114-128) -- brute force over the full 4-index
  tensor expanded from the unique array (test_jk_from_unique_brute_force).
* Rys rule (Roots, Rys.jl:161-174)  -- the Gauss property sum_r w_r x_r^k =
  F_k(X), k < 2n, against 40-digit Boys values (it fixes the n-point rule);
  the reference fits Root1..Root5, whose own last digits stay unpinned here.
* coulomb_rys / all_twoe_ints_rys (Rys.jl:16-151) -- the Rys KATs of
  runtests.jl:130-148, and agreement with the exact HGP values (binary128) and
  THO on random contracted bases up to f (a wrong Recur/Shift term, sign or
  index fails them).
"""
import math

import mpmath
import numpy as np
import pytest

from oracle import oracle as O


def m2ao0(ijk):
    from mi_amd import m2ao
    return m2ao(ijk) - 1


# ------------------------------------------------------------------ Boys
@pytest.mark.parametrize("m,T,val", [(0, 0, 1.0), (0, 1, 0.7468241328124271), (1, 0, 0.33333333333333)])
def test_boys_reference_kats(m, T, val):
    assert O.Fm(m, T) == pytest.approx(val, rel=1.5e-8)


def boys_mp(m, T):
    mpmath.mp.dps = 40
    return mpmath.hyp1f1(m + 0.5, m + 1.5, -T) / (2 * m + 1)


@pytest.mark.parametrize("m", [0, 1, 4, 8, 12])
def test_boys_against_mpmath(m):
    for T in [0.0, 1e-20, 1e-6, 0.3, 1.0, 2.5, 7.0, 12.0, 19.9, 29.99]:
        ref = float(boys_mp(m, T))
        assert O.Fm(m, T) == pytest.approx(ref, rel=1e-14, abs=0)


def test_boys_closed_forms_and_recursion():
    for T in [0.01, 0.5, 3.0, 11.0, 25.0]:
        f0 = 0.5 * math.sqrt(math.pi / T) * math.erf(math.sqrt(T))
        assert O.Fm(0, T) == pytest.approx(f0, rel=1e-14)
        for m in range(6):
            lhs = (2 * m + 1) * O.Fm(m, T)
            rhs = 2 * T * O.Fm(m + 1, T) + math.exp(-T)
            assert lhs == pytest.approx(rhs, rel=1e-13)
    for m in range(10):
        assert O.Fm(m, 0.0) == 1.0 / (2 * m + 1)


def test_farray_branches():
    # T < 30: top order from Fm, downward recursion with e^-T
    F = O.farray(9, 12.5)
    for m in range(9):
        assert F[m] == pytest.approx(float(boys_mp(m, 12.5)), rel=1e-13)
    # T >= 30: the reference's asymptotic branch exactly (no e^-T term)
    T = 41.0
    F = O.farray(6, T)
    mt = 5
    top = math.sqrt(math.pi / 2) * math.prod(range(2 * mt - 1, 0, -2)) / (2 * T) ** (mt + 0.5)
    assert F[mt] == pytest.approx(top, rel=1e-15)
    for m in range(mt - 1, -1, -1):
        assert F[m] == pytest.approx(T * F[m + 1] / (m + 0.5), rel=1e-15)
    assert F[0] == pytest.approx(float(boys_mp(0, T)), rel=1e-14)


@pytest.mark.parametrize("m", [0, 3, 6, 9, 12])
def test_farray_asymptotic_branch_against_mpmath(m):
    """farray's T >= 30 branch (Boys.jl:21-26) for the orders HGP reaches with
    f shells (m <= 12): the reference drops e^{-T}, so it is exact for F_0 up to
    erfc(sqrt T) and off by up to ~1e-4 relative for F_12 at T = 30 (the error
    each order carries is pinned here, against the 40-digit Boys function).
    The binary128 oracle reproduces the branch in its reference mode and is
    exact in its accurate mode."""
    for T in [30.0, 33.0, 45.0, 80.0, 250.0, 1e4]:
        ex = float(boys_mp(m, T))
        ref_branch = math.sqrt(math.pi / 2) * math.prod(range(2 * m - 1, 0, -2)) / (2 * T) ** (m + 0.5)
        F = O.farray(m + 1, T)
        assert F[m] == pytest.approx(ref_branch, rel=1e-15)
        # the branch's own error: the dropped upper incomplete gamma tail
        mpmath.mp.dps = 40
        tail = float(mpmath.gammainc(m + 0.5, T) / (2 * mpmath.mpf(T) ** (m + 0.5)))
        assert F[m] == pytest.approx(ex + tail, rel=4e-16)
        assert O.Fm_quad(m, T, accurate=True) == pytest.approx(ex, rel=1e-15)
        assert O.Fm_quad(m, T, accurate=False) == pytest.approx(ref_branch, rel=1e-15)
    assert abs(O.farray(13, 30.0)[12] / float(boys_mp(12, 30.0)) - 1) > 1e-5


def test_quad_oracle_against_mpmath_tho_fixture():
    """The binary128 oracle against the committed mpmath THO fixture
    (tests/golden/cluster40_outliers.npz, written by make_cluster_outliers.py):
    450 integrals of the 920-function cluster.  Accurate mode: the same value
    to ~1e-16 relative (an independent algorithm, THO instead of HGP, at 40
    digits).  The fixture's hard kinds are where the cheaper restatements miss
    the parity bar, which pins the regime where only the binary128 oracle is
    right: FP64 HGP (kind 0, HRR cancellation), x87 80-bit HGP (kind 1), and
    the reference's own T >= 30 Boys branch (kind 2, literal D5)."""
    import os
    from mi_amd import configs, sub_basis
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cluster40_outliers.npz"))
    wl, _ = configs.cluster_40(0)
    fb = wl.flat()
    qa, qr, d64, d80, idx = [], [], [], [], []
    for k, sh in enumerate(g["shells"]):
        sb = sub_basis(fb, sh)
        sel_i = np.nonzero(g["set_id"] == k)[0]
        sel = g["slot"][sel_i]
        idx.append(sel_i)
        qa.append(O.eri_slots_quad(sb, sel, accurate=True, nthreads=8))
        qr.append(O.eri_slots_quad(sb, sel, accurate=False, nthreads=8))
        d64.append(O.eri_unique_hgp(sb, order=O.ORDER_CHRR, nthreads=8)[sel])
        d80.append(O.eri_unique_hgp_ext(sb, order=O.ORDER_CHRR, nthreads=8)[sel])
    order = np.concatenate(idx)
    ex, kind = g["exact"][order], g["kind"][order]
    qa, qr, d64, d80 = (np.concatenate(x) for x in (qa, qr, d64, d80))
    assert np.all(np.abs(qa - ex) <= 1e-15 * np.abs(ex) + 1e-25)
    bar = lambda x: (np.abs(x - ex) <= 1e-12) | (np.abs(x - ex) <= 1e-10 * np.abs(ex))
    assert not bar(d64)[kind == 0].any()      # FP64 HGP misses every kind-0 integral
    assert (~bar(d80)[kind == 1]).sum() >= 10  # x87 80-bit HGP misses kind-1 integrals
    assert (~bar(qr)[kind == 2]).sum() >= 1    # the reference's T >= 30 branch misses the bar
    assert bar(d64)[kind == 3].all()          # controls: FP64 HGP is fine there


# ------------------------------------------------------------------ AO tables
def test_ao_table_kats():
    assert [O.nao(l) for l in range(3)] == [1, 4, 10]
    assert m2ao0((0, 0, 0)) == 0 and m2ao0((0, 0, 1)) == 3
    lib = O.lib()
    import ctypes
    m = (ctypes.c_int32 * 3)()
    f_order = [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2), (1, 1, 1), (0, 3, 0), (0, 2, 1), (0, 1, 2),
               (0, 0, 3)]
    for k, ijk in enumerate(f_order):
        lib.oracle_ao2m(10 + k, m)
        assert tuple(m) == ijk
    # shift_direction is the FIRST argmax (Julia argmax)
    assert lib.oracle_shift_direction(m2ao0((1, 1, 0))) == 0
    assert lib.oracle_shift_direction(m2ao0((0, 1, 1))) == 1
    assert lib.oracle_shift_direction(m2ao0((1, 2, 1))) == 1


# ------------------------------------------------------------------ VRR / HRR KATs
VRR_KATS = [
    ((0., 0., 0.), (0, 0, 0), (0, 0, 0), 4.37335456733),
    ((0., 0., 0.), (1, 0, 0), (1, 0, 0), 0.182223107579),
    ((0., 0., 0.), (0, 1, 0), (0, 1, 0), 0.182223107579),
    ((0., 0., 0.), (0, 0, 1), (0, 0, 1), 0.182223107579),
    ((0., 0., 0.), (2, 0, 0), (2, 0, 0), 0.223223306785),
    ((0., 0., 0.), (0, 2, 0), (0, 2, 0), 0.223223306785),
    ((0., 0., 0.), (0, 0, 2), (0, 0, 2), 0.223223306785),
    ((1., 2., 3.), (1, 0, 0), (1, 0, 0), -5.63387712455e-06),
    ((1., 2., 3.), (0, 1, 0), (0, 1, 0), -0.000116463120359),
    ((1., 2., 3.), (0, 0, 1), (0, 0, 1), -0.000301178525749),
    ((1., 2., 3.), (2, 0, 0), (2, 0, 0), 0.00022503308545040895),
    ((1., 2., 3.), (0, 2, 0), (0, 2, 0), 0.0006102470883881907),
    ((1., 2., 3.), (0, 0, 2), (0, 0, 2), 0.0013427831014563411),
    ((0., 0., 0.), (1, 1, 0), (1, 1, 0), 0.0136667330685),
    ((0., 0., 0.), (0, 1, 1), (0, 1, 1), 0.0136667330685),
    ((0., 0., 0.), (1, 0, 1), (1, 0, 1), 0.0136667330685),
    ((3., 2., 1.), (1, 1, 0), (1, 1, 0), 5.976771621486971e-5),
    ((3., 2., 1.), (0, 1, 1), (0, 1, 1), 1.5742904443905067e-6),
    ((3., 2., 1.), (1, 0, 1), (1, 0, 1), 4.00292848649699e-6),
]  # test/runtests.jl:159-184

HRR_KATS = [
    ((0., 0., 0.), (0, 0, 0), (0, 0, 0), 0.0136667330685),
    ((0., 0., 0.), (1, 0, 0), (1, 0, 0), 0.00821630976139),
    ((0., 0., 0.), (0, 1, 0), (0, 1, 0), 0.00122024402397),
    ((0., 0., 0.), (0, 0, 1), (0, 0, 1), 0.00821630976139),
    ((0., 0., 0.), (2, 0, 0), (2, 0, 0), 0.0039759617781),
    ((0., 0., 0.), (0, 2, 0), (0, 2, 0), 0.000599953311785),
    ((0., 0., 0.), (0, 0, 2), (0, 0, 2), 0.0039759617781),
    ((1., 2., 3.), (1, 0, 0), (1, 0, 0), -1.1851316496333975e-6),
    ((1., 2., 3.), (0, 1, 0), (0, 1, 0), -4.669991667384835e-6),
    ((1., 2., 3.), (0, 0, 1), (0, 0, 1), -3.474373852654044e-5),
    ((1., 2., 3.), (2, 0, 0), (2, 0, 0), 2.81002247462e-6),
    ((1., 2., 3.), (0, 2, 0), (0, 2, 0), 7.09856891538e-6),
    ((1., 2., 3.), (0, 0, 2), (0, 0, 2), 3.62153023224e-5),
    ((0., 0., 0.), (1, 1, 0), (1, 1, 0), 0.000599953311785),
    ((0., 0., 0.), (0, 1, 1), (0, 1, 1), 0.000599953311785),
    ((0., 0., 0.), (1, 0, 1), (1, 0, 1), 0.0116431617287),
    ((3., 2., 1.), (1, 1, 0), (1, 1, 0), 7.37307761485e-6),
    ((3., 2., 1.), (0, 1, 1), (0, 1, 1), 2.5333243119843164e-7),
    ((3., 2., 1.), (1, 0, 1), (1, 0, 1), 2.452115184675799e-6),
]  # test/runtests.jl:198-222, b = d = (1,0,1)


@pytest.mark.parametrize("literal", [0, O.LIT_D1])
@pytest.mark.parametrize("A,a,c,val", VRR_KATS)
def test_vrr_kats(A, a, c, val, literal):
    # all exponents 1.0 -> the D1 prefactor coincides with the correct one
    v = O.vrr(sum(a), sum(c), 1.0, 1.0, 1.0, 1.0, A, (0, 0, 0), (0, 0, 0), (0, 0, 0), literal)
    assert v[0, m2ao0(a), m2ao0(c)] == pytest.approx(val, rel=1.5e-8)


@pytest.mark.parametrize("A,a,c,val", HRR_KATS)
def test_hrr_kats(A, a, c, val):
    b = d = (1, 0, 1)
    h = O.hrr(sum(a), 2, sum(c), 2, 1.0, 1.0, 1.0, 1.0, A, (0, 0, 0), (0, 0, 0), (0, 0, 0))
    assert h[m2ao0(a), m2ao0(b), m2ao0(c), m2ao0(d)] == pytest.approx(val, rel=1.5e-8)


# ------------------------------------------------------------------ prefactor (D1)
def test_prefactor_closed_form_unequal_exponents():
    """[ss|ss]^(0) for four primitives on one centre is 2 pi^{5/2}/(zeta eta
    sqrt(zeta+eta)) (T = 0, F_0 = 1).  The reference's vrr! (D1) would give
    2 pi^{5/2} (zeta+eta)^{-3/2}."""
    a, b, c, d = 0.5, 0.5, 3.0, 3.0
    z, e = a + b, c + d
    good = 2 * math.pi ** 2.5 / (z * e * math.sqrt(z + e))
    lit = 2 * math.pi ** 2.5 / (z + e) ** 1.5
    O0 = (0.0, 0.0, 0.0)
    assert O.vrr(0, 0, a, b, c, d, O0, O0, O0, O0)[0, 0, 0] == pytest.approx(good, rel=1e-14)
    assert O.vrr(0, 0, a, b, c, d, O0, O0, O0, O0, O.LIT_D1)[0, 0, 0] == pytest.approx(lit, rel=1e-14)
    assert O.coulomb_prim(a, O0, (0, 0, 0), b, O0, (0, 0, 0), c, O0, (0, 0, 0), d, O0, (0, 0, 0)) == \
        pytest.approx(good, rel=1e-14)


def test_eri_tests_from_pyquante():
    """test/runtests.jl:110-127: primitive and distance-scan KATs via THO and HGP."""
    from mi_amd import flat_from_shells
    O0 = (0, 0, 0)
    assert O.coulomb_prim(1, O0, O0, 1, O0, O0, 1, O0, O0, 1, O0, O0) == pytest.approx(4.37335458, rel=1e-8)
    for r, val in [(0.0, 1.1283791670951362), (1.0, 0.8427007900292186), (2.0, 0.49766113257563993),
                   (3.0, 0.33332596983445223), (4.0, 0.2499999961456855), (5.0, 1 / 5), (6.0, 1 / 6),
                   (9.0, 1 / 9)]:
        fb = flat_from_shells([((0, 0, 0), 0, [1.0], [1.0]), ((0, 0, r), 0, [1.0], [1.0])])
        e = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX)
        # (11|22) is iindex slot 4 -> 0-based 3
        assert e[3] == pytest.approx(val, rel=1e-8)
    # (ss|px px) = 0.9403159725793302, (ss|s px) == 0 exactly (runtests.jl:112-113)
    fb = flat_from_shells([((0, 0, 0), 0, [1.0], [1.0]), ((0, 0, 0), 1, [1.0], [1.0])])
    e = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX)
    from mi_amd import iindex
    assert e[iindex(1, 1, 2, 2) - 1] == pytest.approx(0.9403159725793302, rel=1e-12)
    assert e[iindex(1, 1, 1, 2) - 1] == 0.0


# ------------------------------------------------------------------ contracted (D2)
def test_h2_sto3g_contracted_kats():
    from mi_amd import build_basis, flatten, iindex
    from mi_amd.molecules import h2
    fb = flatten(build_basis(h2))
    e = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX)
    kat = {(1, 1, 1, 1): 0.7746059439196398, (2, 2, 2, 2): 0.7746059439196398, (1, 1, 2, 2): 0.5727937653511646,
           (1, 1, 1, 2): 0.4488373301593464, (1, 2, 1, 2): 0.3025451156654606}
    for q, v in kat.items():
        # pyquante2's values carry ~10 digits; the reference checks them at rtol 1.5e-8
        assert e[iindex(*q) - 1] == pytest.approx(v, rel=1e-9)
    lit = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX, literal=O.LIT_D1 | O.LIT_D2 | O.LIT_D3)
    assert abs(lit[0] - kat[(1, 1, 1, 1)]) > 1.0  # the shipped driver is off


# ------------------------------------------------------------------ THO cross-check (D3, VRR/HRR terms)
def random_spec(rng, nshell, lmax=2, kmax=2):
    spec = []
    cents = rng.uniform(-1.5, 1.5, size=(3, 3))
    for _ in range(nshell):
        L = int(rng.integers(0, lmax + 1))
        k = int(rng.integers(1, kmax + 1))
        ex = np.exp(rng.uniform(np.log(0.2), np.log(8.0), size=k))
        spec.append((cents[int(rng.integers(0, 3))], L, ex.tolist(), rng.uniform(0.2, 1, size=k).tolist()))
    return spec


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hgp_equals_tho_on_random_bases(seed):
    from mi_amd import flat_from_shells
    rng = np.random.Generator(np.random.PCG64(seed))
    fb = flat_from_shells(random_spec(rng, 4))
    for order in (O.ORDER_IINDEX, O.ORDER_CHRR):
        h = O.eri_unique_hgp(fb, order=order)
        t = O.eri_unique_tho(fb, order=order)
        assert np.max(np.abs(h - t)) <= 1e-13 * max(1.0, np.max(np.abs(t)))


@pytest.mark.parametrize("seed", [0, 1])
def test_hgp_equals_tho_with_f_shells(seed):
    """f shells (la+lb up to 6, the range the reference's nao table cannot
    reach): HGP VRR/HRR against the independent THO expansion."""
    from mi_amd import flat_from_shells
    rng = np.random.Generator(np.random.PCG64(40 + seed))
    cents = rng.uniform(-1.5, 1.5, size=(3, 3))
    spec = []
    for L in (3, 3, 1):
        k = int(rng.integers(1, 3))
        ex = np.exp(rng.uniform(np.log(0.2), np.log(8.0), size=k))
        spec.append((cents[int(rng.integers(0, 3))], L, ex.tolist(), rng.uniform(0.2, 1, size=k).tolist()))
    fb = flat_from_shells(spec)
    h = O.eri_unique_hgp(fb, order=O.ORDER_CHRR)
    t = O.eri_unique_tho(fb, order=O.ORDER_CHRR)
    assert np.max(np.abs(h - t)) <= 2e-12 * max(1.0, np.max(np.abs(t)))


def test_hgp_equals_tho_ethane_minimal():
    from mi_amd import configs
    fb = configs.ethane_minimal(0).flat()
    h = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX)
    t = O.eri_unique_tho(fb, order=O.ORDER_IINDEX)
    assert np.max(np.abs(h - t)) < 1e-12


def test_literal_d3_changes_p_shells_only():
    from mi_amd import flat_from_shells
    s_only = flat_from_shells([((0, 0, 0), 0, [1.0, 0.3], [0.5, 0.5]), ((0, 0, 1.2), 0, [0.7], [1.0])])
    assert np.array_equal(O.eri_unique_hgp(s_only, literal=O.LIT_D3), O.eri_unique_hgp(s_only))
    with_p = flat_from_shells([((0, 0, 0), 1, [1.0], [1.0]), ((0, 0, 1.2), 0, [0.7], [1.0])])
    assert not np.allclose(O.eri_unique_hgp(with_p, literal=O.LIT_D3), O.eri_unique_hgp(with_p))


# ------------------------------------------------------------------ 8-fold symmetry invariants
def test_full_tensor_symmetry_from_unique():
    """Expanding the unique array to n^4 reproduces all 8 symmetries and
    equals THO element by element (ethane minimal, 16 functions)."""
    from mi_amd import configs
    from mi_amd.indexing import slots
    fb = configs.ethane_minimal(1).flat()
    n = fb.nbf
    e = O.eri_unique_hgp(fb, order=O.ORDER_CHRR)
    i, j, k, l = np.meshgrid(*[np.arange(1, n + 1)] * 4, indexing="ij")
    full = e[slots(n, i.ravel(), j.ravel(), k.ravel(), l.ravel(), 1)].reshape(n, n, n, n)
    assert np.array_equal(full, full.transpose(1, 0, 2, 3))
    assert np.array_equal(full, full.transpose(2, 3, 0, 1))
    t = O.eri_unique_tho(fb, order=O.ORDER_CHRR)
    assert np.max(np.abs(e - t)) < 1e-12


def test_jk_from_unique_brute_force():
    from mi_amd import configs
    from mi_amd.indexing import slots
    fb = configs.ethane_minimal(2).flat()
    n = fb.nbf
    e = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX)
    rng = np.random.Generator(np.random.PCG64(3))
    D = rng.normal(size=(n, n))
    D = D + D.T
    J, K = O.jk_from_unique(n, e, D)
    i, j, k, l = np.meshgrid(*[np.arange(1, n + 1)] * 4, indexing="ij")
    full = e[slots(n, i.ravel(), j.ravel(), k.ravel(), l.ravel(), 0)].reshape(n, n, n, n)
    assert np.allclose(J, np.einsum("ijkl,kl->ij", full, D), rtol=1e-13, atol=1e-12)
    assert np.allclose(K, np.einsum("ikjl,kl->ij", full, D), rtol=1e-13, atol=1e-12)


def test_extended_oracle_arbitrates_cancellation():
    """(p d|p d) with a diffuse p (0.0592) on one centre and a tight d (1206)
    2.1 bohr away: the reference's HRR order cancels ~7 digits in FP64.  The
    exact value, 4.5557863656312922624e-05, was computed once with a 60-digit
    mpmath restatement of the THO formula (src/ERI.jl:15-94) for this quartet of
    the seed-0 split-valence workload (0-based chrr slot 838794)."""
    from mi_amd import configs
    fb = configs.ethane_split_valence(0).flat()
    e64 = O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=8)
    ext = O.eri_unique_hgp_ext(fb, order=O.ORDER_CHRR, nthreads=8)
    exact = 4.5557863656312922624e-05
    assert abs(ext[838794] - exact) < 1e-14
    assert abs(e64[838794] - exact) > 1e-12
    d = np.abs(e64 - ext)
    off = ~((d <= 1e-12) | (d <= 1e-10 * np.abs(ext)))
    assert off.sum() <= 4


# ------------------------------------------------------------------ Rys (src/Rys.jl)
def test_rys_rule_is_the_gauss_rule_of_the_boys_moments():
    """oracle_rys_roots (Roots, Rys.jl:161-174): sum_r w_r x_r^k = F_k(X) for
    k < 2n with x = u/(1+u), against the 40-digit Boys function; nodes in (0,1)."""
    for n in range(1, 11):
        for X in [0.0, 1e-9, 0.3, 5.0, 29.0, 31.0, 60.0, 150.0]:
            u, w = O.rys_roots(n, X)
            x = u / (1 + u)
            assert np.all((x > 0) & (x < 1)) and np.all(np.diff(x) > 0)
            for k in range(2 * n):
                Fk = float(boys_mp(k, X))
                assert float(np.sum(w * x ** k)) == pytest.approx(Fk, rel=5e-14 if n <= 7 else 1e-12), (n, X, k)


def test_rys_reference_kats():
    """test/runtests.jl:130-148 through the Rys restatement (all_twoe_ints_rys
    on normalised primitives): (ss|ss) at one centre, (ss|px px), the
    symmetry zero (ss|s px) == 0, and the distance scan."""
    from mi_amd import flat_from_shells, iindex
    for r, val in [(0.0, 1.1283791670951362), (1.0, 0.8427007900292186), (2.0, 0.49766113257563993),
                   (3.0, 0.33332596983445223), (4.0, 0.2499999961456855), (5.0, 1 / 5), (6.0, 1 / 6),
                   (7.0, 1 / 7), (8.0, 1 / 8), (9.0, 1 / 9)]:
        fb = flat_from_shells([((0, 0, 0), 0, [1.0], [1.0]), ((0, 0, r), 0, [1.0], [1.0])])
        e = O.eri_unique_rys(fb, order=O.ORDER_IINDEX)
        assert e[iindex(1, 1, 2, 2) - 1] == pytest.approx(val, rel=1e-8)
        assert e[0] == pytest.approx(1.128379167, rel=1e-9)  # coulomb_rys(c,c,c,c)
    fb = flat_from_shells([((0, 0, 0), 0, [1.0], [1.0]), ((0, 0, 0), 1, [1.0], [1.0])])
    e = O.eri_unique_rys(fb, order=O.ORDER_IINDEX)
    assert e[iindex(1, 1, 2, 2) - 1] == pytest.approx(0.9403159725793302, rel=1e-12)
    assert e[iindex(1, 1, 1, 2) - 1] == 0.0


@pytest.mark.parametrize("lmax,seed", [(2, 3), (3, 4)])
def test_rys_equals_hgp_and_tho(lmax, seed):
    """Rys quadrature and HGP recurrences are different algorithms for the
    same integrals: the Rys restatement (FP64) against the exact value of the
    HGP algorithm (binary128, accurate Boys) on random contracted bases up to
    f, and against THO (runtests.jl:60 compares all_twoe_ints with
    all_twoe_ints_rys at rtol 1e-7)."""
    from mi_amd import flat_from_shells
    rng = np.random.default_rng(seed)
    fb = flat_from_shells(random_spec(rng, 5, lmax=lmax, kmax=2))
    rys = O.eri_unique_rys(fb, order=O.ORDER_CHRR)
    ex = O.eri_unique_hgp_quad(fb, order=O.ORDER_CHRR, accurate=True)
    d = np.abs(rys - ex)
    assert np.all((d <= 1e-12) | (d <= 1e-10 * np.abs(ex))), float(d.max())
    assert float(d.max()) <= 1e-13 * max(1.0, float(np.abs(ex).max()))
    if lmax <= 2:
        tho = O.eri_unique_tho(fb, order=O.ORDER_CHRR)
        assert np.allclose(rys, tho, rtol=1e-10, atol=1e-12)
