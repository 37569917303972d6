"""Host-side mirror of the reference basis model, workload generators and the
C-ABI library (CPU only: the library must load and export every declared
symbol; no compute call needs a GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_basis_kats():
    """test/runtests.jl:27-47."""
    from mi_amd import addbf, build_basis, cgbf, nao, overlap, pgbf
    from mi_amd.molecules import h2
    s = pgbf(1.0)
    px = pgbf(1.0, 0, 0, 0, 1, 0, 0)
    assert s(0, 0, 0) == pytest.approx(0.71270547, rel=1e-8)
    assert s.norm == pytest.approx(0.7127054703549901, rel=1e-15)
    assert px.norm == pytest.approx(1.4254109407099802, rel=1e-15)
    assert px(0, 0, 0) == 0
    c = cgbf(0.0, 0.0, 0.0)
    addbf(c, 1, 1)
    assert c(0, 0, 0) == pytest.approx(0.71270547, rel=1e-8)
    c2 = cgbf(0, 0, 0)
    addbf(c2, 1, 0.2)
    addbf(c2, 0.5, 0.2)
    assert overlap(c2, c2) == pytest.approx(1.0, rel=1e-14)
    assert [nao(l) for l in range(3)] == [1, 4, 10]
    bfs = build_basis(h2)
    assert len(bfs) == 2
    l, r = bfs
    assert overlap(l, l) == pytest.approx(1, rel=1e-14)
    assert overlap(l, r) == pytest.approx(0.6647387449282997, rel=1e-12)


def test_norms_match_oracle():
    from mi_amd import pgbf
    for e in (0.1, 1.0, 37.0):
        for ijk in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 0, 0), (1, 2, 0), (1, 1, 1)]:
            assert pgbf(e, 0, 0, 0, *ijk).norm == pytest.approx(O.pgbf_norm(e, *ijk), rel=1e-14)


def test_flat_from_shells_equals_flatten():
    from mi_amd import basis_from_shells, flat_from_shells, flatten
    from mi_amd import configs
    spec = configs.ethane_split_valence(0).spec
    a = flat_from_shells(spec)
    b = flatten(basis_from_shells(spec))
    for x, y in zip(a.arrays(), b.arrays()):
        if x.dtype == np.float64:
            assert np.allclose(x, y, rtol=1e-13, atol=0)
        else:
            assert np.array_equal(x, y)


def test_workload_sizes():
    from mi_amd import configs
    assert configs.ethane_minimal(0).nbf == 16
    assert O.n_unique(16) == 9316
    assert configs.ethane_split_valence(0).nbf == 60
    assert O.n_unique(60) == 1675365
    wl, D = configs.cluster_40(0)
    assert wl.nbf == 920 and D.shape == (920, 920) and np.allclose(D, D.T)
    mols = configs.batch_molecules(0, nmol=8)
    assert all(3 <= len(m.heavy) <= 24 for m in mols)
    for m in mols:
        xyz = np.array([s[0] for s in m.spec])
        d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1)
        d = d[d > 1e-9]
        assert d.min() * 0.52918 >= 0.9 - 1e-9


def test_workloads_are_seeded():
    from mi_amd import configs
    a = configs.ethane_split_valence(3).flat().arrays()
    b = configs.ethane_split_valence(3).flat().arrays()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_library_exports_every_declared_symbol():
    from mi_amd import _lib
    L = _lib.load()
    with open(os.path.join(ROOT, "include", "mi_hgp.h")) as f:
        hdr = f.read()
    declared = set(re.findall(r"\b(mi_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(_lib.EXPORTED), declared ^ set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    assert L.mi_n_unique(60) == 1675365
    assert L.mi_lmax() >= 2


def test_library_reports_missing_device_as_error():
    """No GPU in the CPU suite: mi_create must fail cleanly (no fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from mi_amd import _lib
    h = ctypes.c_void_p()
    assert _lib.load().mi_create(0, ctypes.byref(h)) == _lib.MI_EDEVICE


def test_flop_model_matches_published_table():
    from mi_amd.flops import vrr_ops
    table = {(0, 0): 1, (1, 0): 11, (1, 1): 72, (2, 0): 51, (2, 1): 256, (2, 2): 920, (3, 1): 668, (3, 2): 2244,
             (3, 3): 5920, (4, 2): 4585, (4, 3): 11841, (4, 4): 24887}
    for k, v in table.items():
        assert vrr_ops(*k) == v


def test_sub_basis_integrals_equal_the_full_basis():
    """Every unique integral of sub_basis(fb, shells) equals the integral of
    the same four functions in fb (oracle, iindex order, both arrays): the
    helper keeps each shell's primitives, norms and function scales."""
    from mi_amd import flat_from_shells, iindex, shell_functions, sub_basis
    rng = np.random.Generator(np.random.PCG64(5))
    spec = []
    for s in range(7):
        L = int(rng.integers(0, 3))
        k = int(rng.integers(1, 4))
        e = np.sort(np.exp(rng.uniform(np.log(0.2), np.log(20.0), k)))[::-1]
        spec.append((rng.normal(0.0, 1.5, 3), L, e.tolist(), rng.uniform(0.1, 1.0, k).tolist()))
    fb = flat_from_shells(spec)
    full = O.eri_unique_hgp(fb, order=O.ORDER_IINDEX)
    shells = [5, 1, 3, 6]
    sb = sub_basis(fb, shells)
    fmap = shell_functions(fb, shells) + 1  # 1-based full-basis index of each sub-basis function
    assert sb.nbf == fmap.size
    sub = O.eri_unique_hgp(sb, order=O.ORDER_IINDEX)
    n = sb.nbf
    got, ref = [], []
    for i in range(1, n + 1):
        for j in range(1, i + 1):
            for k in range(1, n + 1):
                for l in range(1, k + 1):
                    if k * (k - 1) // 2 + l > i * (i - 1) // 2 + j:  # each pair of pairs once (ij >= kl)
                        continue
                    got.append(sub[iindex(i, j, k, l) - 1])
                    ref.append(full[iindex(int(fmap[i - 1]), int(fmap[j - 1]), int(fmap[k - 1]),
                                           int(fmap[l - 1])) - 1])
    got, ref = np.array(got), np.array(ref)
    assert got.size == O.n_unique(n)
    d = np.abs(got - ref)
    assert np.all((d <= 1e-12) | (d <= 1e-10 * np.abs(ref))), d.max()
    assert np.abs(ref).max() > 1e-2  # non-trivial values


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and says why
    (before touching the GPU or the library)."""
    import json
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_python_api_4d_expansion_matches_the_reference_loop():
    """MolecularIntegrals.all_twoe_ints_4d expands the iindex array with the
    0-based index algebra of python/molecular_integrals.py:467-482; the
    vectorised expansion equals that loop, element for element."""
    from mi_amd.molecular_integrals import _expand_4d
    n = 5
    ints = np.random.Generator(np.random.PCG64(0)).normal(size=n * (n + 1) * (n * n + n + 2) // 8)

    def triangle(i):
        return i * (i + 1) // 2

    def iindex(i, j, k, l):
        ij = triangle(max(i, j)) + min(i, j)
        kl = triangle(max(k, l)) + min(k, l)
        return triangle(max(ij, kl)) + min(ij, kl)

    ref = np.zeros((n, n, n, n))
    for i in range(n):
        for j in range(n):
            for k in range(n):
                for l in range(n):
                    ref[i, j, k, l] = ints[iindex(i, j, k, l)]
    assert np.array_equal(_expand_4d(ints, n), ref)


def test_python_api_names_and_errors():
    """Atom / Basis / MolecularIntegrals of python/molecular_integrals.py:
    Atom xyz in bohr, Basis(atoms, name) through build_basis (names matched
    without case), unknown methods ValueError -- all before any device call."""
    from mi_amd.molecular_integrals import Atom, Basis, MolecularIntegrals
    atoms = [Atom(1, (0.0, 0.0, 0.0)), Atom(1, (1.4, 0.0, 0.0))]
    assert atoms[1].xyz == (1.4, 0.0, 0.0) and atoms[0].atno == 1
    b = Basis(atoms, "STO3G")
    assert len(b) == 2 and b.name == "sto3g"
    assert len(Basis([Atom(6, (0, 0, 0))], "cc-pVDZ")) == 15
    with pytest.raises(KeyError):
        Basis(atoms, "nonsense")
    mi = MolecularIntegrals()
    with pytest.raises(ValueError):
        mi.all_twoe_ints(b, method="bogus")


def test_rys_kernel_rule_is_the_gauss_rule():
    """The Rys kernel's rule (host copy of its tables, mi_rys_rule) is the
    Gauss rule of e^{-T x} x^{-1/2}/2 on [0, 1]: sum_r w_r x_r^k = F_k(T) for
    k < 2n (the moments determine the n-point rule uniquely), checked against
    the oracle's binary128 Boys function on table intervals, interval edges
    and the large-T limit; and it agrees root by root with the oracle's rule
    (computed by a different method: Golub-Welsch with Jacobi rotations)."""
    from mi_amd import _lib
    from oracle import oracle as O
    L = _lib.load()
    nmax = int(L.mi_rys_nmax())
    assert nmax >= 5  # the reference's Roots covers n <= 5 (src/Rys.jl:161-174)
    rng = np.random.default_rng(5)
    Ts = [0.0, 1e-12, 1e-3, 0.5, 1.0, 2.0 - 1e-13, 2.0, 9.99, 29.5, 30.0, 41.0, 47.7, 55.2, 68.0, 79.9, 80.0, 120.0,
          1e3, 1e6] + list(rng.uniform(0, 90, 40))
    worst = 0.0
    for n in range(1, nmax + 1):
        for T in Ts:
            x, w = _lib.rys_rule(n, T)
            assert np.all((x > 0) & (x < 1)) and np.all(w > 0) and np.all(np.diff(x) > 0), (n, T)
            for k in range(2 * n):
                Fk = O.Fm_quad(k, T, accurate=True)
                rel = abs(float(np.sum(w * x ** k)) - Fk) / Fk
                worst = max(worst, rel)
                assert rel < 2e-14, (n, T, k, rel)
            u, wo = O.rys_roots(n, T)
            xo = u / (1 + u)
            assert np.allclose(x, xo, rtol=5e-15, atol=0), (n, T, x, xo)
            assert np.allclose(w, wo, rtol=5e-15, atol=0), (n, T, w, wo)
    print(f"worst moment residual {worst:.2e}")
    with pytest.raises(_lib.MiError):
        _lib.rys_rule(nmax + 1, 1.0)
    with pytest.raises(_lib.MiError):
        _lib.rys_rule(1, -1.0)


def test_tuning_file_rows_are_well_formed():
    """hgp_tuning.txt (read by mi_create; written by mi_amd/autotune.py):
    '#' comment lines, then rows 'class bra_bucket ket_bucket variant G
    regime' with in-range fields, one row per key, register-kernel rows only
    for classes that have a generated register kernel, lane and wavefront
    counts each kernel family accepts, and every canonical class tuned for the
    small- and large-launch regimes."""
    import importlib.util
    from mi_amd.hgp_tables import quartet_classes
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location(
        "gen_hgp", os.path.join(root, "molecularintegrals.jl_amd", "csrc", "gen_hgp.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    classes = quartet_classes()
    reg = {k for k, c in enumerate(classes) if gen.is_register_class(c)}
    allowed = {0: {1, 2, 4}, 1: {1, 2, 4, 8, 16, 32, 64}, 2: {1, 2, 4, 8}, 3: {1, 2, 4}, 4: {1},
               5: {1, 2, 4}, 6: {16, 32, 64}}
    keys = set()
    path = os.path.join(root, "molecularintegrals.jl_amd", "hgp_tuning.txt")
    with open(path) as f:
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            c, bb, kb, var, G, regime = (int(x) for x in line.split())
            assert 0 <= c < len(classes) and 0 <= bb < 5 and 0 <= kb < 5 and regime in (0, 1, 2), line
            assert var in allowed and G in allowed[var], line
            assert var != 1 or c in reg, line
            assert (c, bb, kb, regime) not in keys, line
            keys.add((c, bb, kb, regime))
    for regime in (0, 1):
        for c in range(55):  # the canonical classes (L_A >= L_B, L_C >= L_D)
            assert all((c, bb, kb, regime) in keys for bb in range(5) for kb in range(5)), (classes[c], regime)
