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
