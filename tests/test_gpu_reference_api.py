"""The reference's own entry points and timing workloads through the GPU:
the Python API (python/molecular_integrals.py:346-482) routed to the C ABI,
and ethane in the named basis sets of test/timing.jl:81-89 (6-31G, 30
functions; cc-pVDZ, 60 Cartesian functions), whose HGP timings README.md:50-59
publishes.  Bar: count and order exact, every integral within 1e-12 absolute
or 1e-10 relative of the oracle (binary128 arbitration where the FP64 oracle
cancels, as in test_gpu_parity)."""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import assert_eri_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["default", "hgp"])
def test_python_entry_point_runs_hgp_on_the_gpu(method):
    from mi_amd import flatten
    from mi_amd.molecular_integrals import Atom, Basis, MolecularIntegrals
    from mi_amd.molecules import ethane
    atoms = [Atom(a.atno, a.xyz) for a in ethane]
    b = Basis(atoms, "6-31G")
    mi = MolecularIntegrals()
    got = mi.all_twoe_ints(b, method=method)
    fb = flatten(b._basis)
    ref = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX, nthreads=8)
    assert got.size == 108345
    assert_eri_close(got, ref, f"python API method={method}", fb, O.ORDER_IINDEX)
    g4 = mi.all_twoe_ints_4d(b, method=method)
    assert g4.shape == (30, 30, 30, 30)
    assert np.array_equal(g4, g4.transpose(1, 0, 2, 3)) and np.array_equal(g4, g4.transpose(2, 3, 0, 1))
    assert g4[3, 1, 7, 2] == got[O.lib().oracle_slot(30, 4, 2, 8, 3, O.ORDER_IINDEX)]


def test_python_entry_point_h2_known_answers():
    """test/runtests.jl:48-52 values through MolecularIntegrals().all_twoe_ints_4d."""
    from mi_amd.molecular_integrals import Atom, Basis, MolecularIntegrals
    from mi_amd.molecules import h2
    g4 = MolecularIntegrals().all_twoe_ints_4d(Basis([Atom(a.atno, a.xyz) for a in h2], "sto3g"), method="hgp")
    assert g4[0, 0, 0, 0] == pytest.approx(0.7746059439196398, rel=1e-9)
    assert g4[0, 0, 1, 1] == pytest.approx(0.5727937653511646, rel=1e-9)
    assert g4[0, 0, 0, 1] == pytest.approx(0.4488373301593464, rel=1e-9)
    assert g4[0, 1, 0, 1] == pytest.approx(0.3025451156654606, rel=1e-9)


@pytest.mark.parametrize("name,nbf", [("6-31G", 30), ("cc-pvdz", 60)])
@pytest.mark.parametrize("order", [O.ORDER_CHRR, O.ORDER_IINDEX])
def test_ethane_named_basis_timing_workloads(ctx, name, nbf, order):
    from mi_amd import build_basis, flatten
    from mi_amd.molecules import ethane
    fb = flatten(build_basis(ethane, name))
    assert fb.nbf == nbf
    got = ctx.eri_unique(fb, order=order)
    ref = O.eri_unique_hgp(fb, order=order, nthreads=8)
    assert got.size == nbf * (nbf + 1) * (nbf * nbf + nbf + 2) // 8
    assert_eri_close(got, ref, f"ethane/{name} order={order}", fb, order)
