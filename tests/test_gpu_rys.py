"""GPU parity of the Rys quadrature path (MI_METHOD_RYS; all_twoe_ints_rys,
src/Rys.jl:16-25) through the C ABI.

Oracle: the FP64 restatement of coulomb_rys (oracle_eri_unique_rys, pinned in
test_oracle.py by the Gauss property of its rule, the reference's Rys KATs and
agreement with the exact HGP values) on small bases; on larger ones the FP64
HGP oracle (the same integrals by another algorithm), with the binary128 HGP
value deciding where the FP64 oracle is itself off (assert_eri_close).  Bar:
count and order exact, every integral within 1e-12 absolute or 1e-10 relative.
"""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import assert_eri_close, random_spec, within

pytestmark = pytest.mark.gpu

RYS = 1  # MI_METHOD_RYS


def test_rys_reference_kats(ctx):
    """test/runtests.jl:130-148 (coulomb_rys) through the GPU Rys path."""
    from mi_amd import flat_from_shells, iindex
    for r, val in [(0.0, 1.1283791670951362), (1.0, 0.8427007900292186), (2.0, 0.49766113257563993),
                   (3.0, 0.33332596983445223), (4.0, 0.2499999961456855), (5.0, 1 / 5), (6.0, 1 / 6),
                   (7.0, 1 / 7), (8.0, 1 / 8), (9.0, 1 / 9)]:
        fb = flat_from_shells([((0, 0, 0), 0, [1.0], [1.0]), ((0, 0, r), 0, [1.0], [1.0])])
        e = ctx.eri_unique(fb, order=O.ORDER_IINDEX, method=RYS)
        assert e[iindex(1, 1, 2, 2) - 1] == pytest.approx(val, rel=1e-8)
        assert e[0] == pytest.approx(1.128379167, rel=1e-9)
    fb = flat_from_shells([((0, 0, 0), 0, [1.0], [1.0]), ((0, 0, 0), 1, [1.0], [1.0])])
    e = ctx.eri_unique(fb, order=O.ORDER_IINDEX, method=RYS)
    assert e[iindex(1, 1, 2, 2) - 1] == pytest.approx(0.9403159725793302, rel=1e-12)
    assert e[iindex(1, 1, 1, 2) - 1] == 0.0


@pytest.mark.parametrize("order", [O.ORDER_IINDEX, O.ORDER_CHRR])
@pytest.mark.parametrize("seed", [0, 1])
def test_rys_matches_the_rys_oracle(ctx, seed, order):
    """Ethane-like 16-function workload: every integral against the Rys
    restatement of the reference algorithm."""
    from mi_amd import configs
    fb = configs.ethane_minimal(seed).flat()
    got = ctx.eri_unique(fb, order=order, method=RYS)
    ref = O.eri_unique_rys(fb, order=order, nthreads=8)
    assert got.size == 9316
    assert_eri_close(got, ref, f"rys ethane_minimal seed={seed} order={order}", fb, order)


@pytest.mark.parametrize("lmax,seed", [(2, 0), (2, 1), (3, 2), (3, 3)])
def test_rys_random_bases_every_class(ctx, lmax, seed):
    """Random contracted bases up to f (3-7 roots): the GPU Rys path against
    the Rys oracle, and against the GPU HGP path (different algorithms, the
    same integrals)."""
    from mi_amd import flat_from_shells
    rng = np.random.Generator(np.random.PCG64(700 + seed))
    fb = flat_from_shells(random_spec(rng, 6, lmax=lmax, kmax=2))
    got = ctx.eri_unique(fb, order=O.ORDER_CHRR, method=RYS)
    ref = O.eri_unique_rys(fb, order=O.ORDER_CHRR, nthreads=8)
    assert_eri_close(got, ref, f"rys random lmax={lmax} seed={seed}", fb, O.ORDER_CHRR)
    hgp = ctx.eri_unique(fb, order=O.ORDER_CHRR)
    assert within(got, hgp).all(), float(np.abs(got - hgp).max())


def test_rys_split_valence_full_array(ctx):
    """The 60-function split-valence workload (d shells, deep s contractions):
    all 1,675,365 integrals against the FP64 HGP oracle."""
    from mi_amd import configs
    fb = configs.ethane_split_valence(0).flat()
    got = ctx.eri_unique(fb, order=O.ORDER_CHRR, method=RYS)
    ref = O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=8)
    assert_eri_close(got, ref, "rys ethane_split_valence", fb, O.ORDER_CHRR)


def test_rys_reference_api_and_timing_workloads(ctx):
    """MolecularIntegrals().all_twoe_ints(basis, method='rys')
    (python/molecular_integrals.py:409-440) on the reference's timing
    workloads (ethane STO-3G and 6-31G, test/timing.jl:81-89), against the
    Rys oracle; cc-pVDZ (d shells) against the FP64 HGP oracle."""
    from mi_amd import flatten
    from mi_amd.molecular_integrals import Atom, Basis, MolecularIntegrals
    from mi_amd.molecules import ethane
    atoms = [Atom(a.atno, a.xyz) for a in ethane]
    mi = MolecularIntegrals()
    for name in ("sto3g", "6-31G", "cc-pvdz"):
        b = Basis(atoms, name)
        got = mi.all_twoe_ints(b, method="rys")
        fb = flatten(b._basis)
        if name == "cc-pvdz":
            ref = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX, nthreads=8)
        else:
            ref = O.eri_unique_rys(fb, order=O.ORDER_IINDEX, nthreads=8)
        assert_eri_close(got, ref, f"rys ethane/{name}", fb, O.ORDER_IINDEX)


def test_rys_screened_batched_and_jk(ctx):
    """The Rys kernel in the other execution forms: a screened batched plan
    (Schwarz pass and screening as for HGP) against per-molecule plain
    outputs, and a J/K build against the oracle algebra."""
    from mi_amd import flat_from_shells
    rng = np.random.Generator(np.random.PCG64(91))
    fbs = [flat_from_shells(random_spec(rng, 5 + k, lmax=2, kmax=3)) for k in range(3)]
    import torch
    plan = ctx.batch_plan(fbs, order=O.ORDER_CHRR, schwarz_cut=1e-12, method=RYS)
    out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
    plan.execute(out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    offs = plan.mol_offsets()
    for m, fb in enumerate(fbs):
        ref = O.eri_unique_hgp(fb, order=O.ORDER_CHRR)
        seg = got[offs[m]:offs[m + 1]]
        d = np.abs(seg - ref)
        assert np.all((d <= 1e-12) | (d <= 1e-10 * np.abs(ref)) | (d <= 1.001e-12)), (m, float(d.max()))
    plan.close()
    fb = fbs[2]
    n = fb.nbf
    D = rng.standard_normal((n, n))
    D = (D + D.T) / (2 * n)
    J, K = ctx.jk(fb, D, method=RYS)
    Jr, Kr = O.jk_from_unique(n, O.eri_unique_hgp(fb, order=O.ORDER_IINDEX), D)
    assert np.allclose(J, Jr, rtol=1e-11, atol=1e-12) and np.allclose(K, Kr, rtol=1e-11, atol=1e-12)


def _ryg_rows(ctx, lg):
    """Tuning rows that put every class the Rys lane-group kernel (variant 6)
    can take on it, with lg lanes per quartet, in every regime."""
    from mi_amd.hgp_tables import quartet_classes
    rows = []
    for k in range(len(quartet_classes())):
        try:
            ctx.set_tuning([(k, 0, 0, 6, lg)])
        except Exception:
            continue
        rows += [(k, bb, kb, 6, lg, -1) for bb in range(5) for kb in range(5)]
    ctx.set_tuning([])
    return rows


@pytest.mark.parametrize("lg", [16, 32, 64])
def test_rys_lane_group_kernel(lg):
    """Variant 6 (groups of lg lanes per shell quartet, several quartets per
    wavefront) on random contracted bases up to f: unique array in HGP and Rys
    plans against the oracle, a screened batch, and J/K."""
    import torch
    from mi_amd import Context, flat_from_shells
    c = Context(0)
    rows = _ryg_rows(c, lg)
    assert len(rows) > 25 * 40  # most classes up to f fit
    c.set_tuning(rows)
    rng = np.random.Generator(np.random.PCG64(800 + lg))
    fb = flat_from_shells(random_spec(rng, 7, lmax=3, kmax=3))
    ref = O.eri_unique_hgp(fb, order=O.ORDER_CHRR, nthreads=8)
    for method in (0, RYS):
        plan = c.plan(fb, order=O.ORDER_CHRR, method=method)
        assert any(l[4].startswith("Z") for l in plan.launches())
        plan.close()
        got = c.eri_unique(fb, order=O.ORDER_CHRR, method=method)
        assert_eri_close(got, ref, f"lane-group lg={lg} method={method}", fb, O.ORDER_CHRR)
    fbs = [flat_from_shells(random_spec(rng, 5 + k, lmax=2, kmax=3)) for k in range(3)]
    plan = c.batch_plan(fbs, order=O.ORDER_CHRR, schwarz_cut=1e-12)
    out = torch.empty(plan.n_unique, dtype=torch.float64, device="cuda")
    plan.execute(out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    offs = plan.mol_offsets()
    for m, f in enumerate(fbs):
        r = O.eri_unique_hgp(f, order=O.ORDER_CHRR)
        d = np.abs(got[offs[m]:offs[m + 1]] - r)
        assert np.all((d <= 1e-12) | (d <= 1e-10 * np.abs(r)) | (d <= 1.001e-12)), (m, float(d.max()))
    plan.close()
    n = fb.nbf
    D = rng.standard_normal((n, n))
    D = (D + D.T) / (2 * n)
    J, K = c.jk(fb, D)
    Jr, Kr = O.jk_from_unique(n, O.eri_unique_hgp(fb, order=O.ORDER_IINDEX, nthreads=8), D)
    assert np.allclose(J, Jr, rtol=1e-11, atol=1e-12) and np.allclose(K, Kr, rtol=1e-11, atol=1e-12)
    c.close()
