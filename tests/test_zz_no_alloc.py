"""mi_eri_unique_noalloc (the entry point that allocates no device memory)
against mi_eri_unique, through the C ABI: plain FP64 bit for bit (same plan
and kernels, one stream) and the mixed screened form at the parity bar (its
Schwarz bounds come from a pass inside the execute), on two calls
in a row, the second with a smaller basis (the static scratch area keeps the
first call's contents, which the second must not depend on); a basis beyond
the area is refused with MI_ENOMEM and the next call still works."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(ctx, fb, order=1, precision=0, tau=0.0, cut=0.0):
    import torch
    from mi_amd import _lib
    args = _lib.BasisArgs(fb)
    n = int(ctx.L.mi_n_unique(fb.nbf))
    out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    o = _lib.opts(order, precision, tau, cut)
    o.stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = ctx.L.mi_eri_unique_noalloc(ctx.h, ctypes.byref(args.struct), ctypes.byref(o),
                                     ctypes.c_void_p(out.data_ptr()), n)
    return rc, out.cpu().numpy()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_noalloc_matches_mi_eri_unique_twice(ctx, mode):
    from mi_amd import configs, flat_from_shells
    from mi_amd import _lib
    kw = {} if mode == "fp64" else dict(precision=1, tau=1e-8, schwarz_cut=1e-12)
    rng = np.random.Generator(np.random.PCG64(5))
    big = configs.ethane_split_valence(2).flat()                         # 60 functions, d shells
    f_spec = [(rng.uniform(-2, 2, 3), L, [float(np.exp(rng.uniform(-2, 5)))], [1.0]) for L in (3, 2, 1, 0, 3, 1)]
    small = flat_from_shells(f_spec)                                     # 34 functions, f shells
    for fb, order in ((big, 1), (small, 0)):
        ref = ctx.eri_unique(fb, order=order, **kw)
        rc, got = _noalloc(ctx, fb, order, kw.get("precision", 0), kw.get("tau", 0.0), kw.get("schwarz_cut", 0.0))
        assert rc == _lib.MI_OK, ctx.L.mi_last_error(ctx.h)
        assert got.shape == ref.shape
        if mode == "fp64":  # same plan and kernels: the same bits
            assert np.array_equal(got.view(np.int64), ref.view(np.int64)), (fb.nbf, float(np.nanmax(np.abs(got - ref))))
        else:  # screening bounds computed per execute instead of at plan build: the parity bar
            d = np.abs(got - ref)
            assert np.all((d <= 1e-12) | (d <= 1e-10 * np.abs(ref))), (fb.nbf, float(np.nanmax(d)))


def test_noalloc_refuses_a_basis_beyond_the_scratch_area(ctx):
    """30 s shells of 48 primitives each on three close centres: 465 shell
    pairs of 2,304 kept primitive pairs, 1.07M primitive-pair records (64 B
    each, 69 MB) -- beyond the 64 MiB area, while the output (30 functions,
    108,345 integrals) is small.  The call must refuse with MI_ENOMEM and a
    message; the output buffer passed has the full length, so even a call
    that did run would stay inside it.  The next call still works."""
    import torch
    from mi_amd import configs, flat_from_shells, _lib
    rng = np.random.Generator(np.random.PCG64(8))
    cents = [np.zeros(3), np.array([0.0, 0.0, 1.2]), np.array([0.0, 1.1, 0.4])]
    spec = [(cents[k % 3], 0, list(np.sort(np.exp(rng.uniform(np.log(0.5), np.log(4.0), 48)))[::-1]),
             list(rng.uniform(0.1, 1.0, 48))) for k in range(30)]
    fb = flat_from_shells(spec)
    args = _lib.BasisArgs(fb)
    o = _lib.opts(1, 0, 0.0, 0.0)
    n = int(ctx.L.mi_n_unique(fb.nbf))
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    rc = ctx.L.mi_eri_unique_noalloc(ctx.h, ctypes.byref(args.struct), ctypes.byref(o),
                                     ctypes.c_void_p(out.data_ptr()), n)
    assert rc == _lib.MI_ENOMEM, (rc, ctx.L.mi_last_error(ctx.h))
    assert b"static scratch area" in ctx.L.mi_last_error(ctx.h)
    small = configs.ethane_minimal(0).flat()
    rc, got = _noalloc(ctx, small)
    assert rc == _lib.MI_OK and np.array_equal(got, ctx.eri_unique(small, order=1))
