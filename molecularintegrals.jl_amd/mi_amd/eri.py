"""Drop-in replacements for the reference's ERI drivers, running on the GPU.

* ``all_twoe_ints_chrr(bfs)`` -- rpmuller/MolecularIntegrals.jl src/HGP.jl:22-34.
  Same argument (a ``Basis``), same return (1-d float64 array of length
  n(n+1)(n^2+n+2)/8), same slot order (enumeration of ``iiterator``,
  src/Basis.jl:203).  Values are the normalised contracted integrals
  ``coulomb(bfs[i],bfs[j],bfs[k],bfs[l])`` (src/ERI.jl:102).
* ``all_twoe_ints(bfs)`` -- src/ERI.jl:104-112, slot = ``iindex(i,j,k,l)``.

Both call the C ABI (include/mi_hgp.h) through ``Context``; errors surface as
``MiError`` (the reference raises Julia exceptions, e.g. a BoundsError for
L beyond its tables).
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _lib
from .basis import Basis, FlatBasis, flatten


class Context:
    """One device context (``mi_create``).  Not shared across threads."""

    def __init__(self, device: int = 0):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        rc = self.L.mi_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise _lib.MiError(rc, f"mi_create(device={device}) failed (no usable GPU?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.mi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            raise _lib.MiError(rc, self.L.mi_last_error(self.h).decode())

    def eri_unique(self, fb: FlatBasis, order=_lib.MI_ORDER_CHRR, precision=_lib.MI_PREC_FP64, tau=0.0,
                   schwarz_cut=0.0, method=_lib.MI_METHOD_HGP) -> np.ndarray:
        """All unique ERIs (host array); method MI_METHOD_RYS computes them by
        Rys quadrature (all_twoe_ints_rys, src/Rys.jl:16-25)."""
        args = _lib.BasisArgs(fb)
        n = int(self.L.mi_n_unique(fb.nbf))
        out = np.empty(n, dtype=np.float64)
        o = _lib.opts(order, precision, tau, schwarz_cut, method=method)
        self.check(self.L.mi_eri_unique(self.h, ctypes.byref(args.struct), ctypes.byref(o),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out

    def jk(self, fb: FlatBasis, D, precision=_lib.MI_PREC_FP64, tau=0.0, schwarz_cut=0.0,
           method=_lib.MI_METHOD_HGP):
        """Direct Coulomb and exchange (host arrays): J_ij = sum (ij|kl) D_kl,
        K_ij = sum (ik|jl) D_kl; make2JmK's G is 2J - K (src/ERI.jl:114-128)."""
        args = _lib.BasisArgs(fb)
        n = fb.nbf
        D = np.ascontiguousarray(D, dtype=np.float64).reshape(n, n)
        J = np.empty((n, n))
        K = np.empty((n, n))
        o = _lib.opts(_lib.MI_ORDER_JK, precision, tau, schwarz_cut, method=method)
        dp = ctypes.POINTER(ctypes.c_double)
        self.check(self.L.mi_jk(self.h, ctypes.byref(args.struct), ctypes.byref(o), D.ctypes.data_as(dp),
                                J.ctypes.data_as(dp), K.ctypes.data_as(dp)))
        return J, K

    def set_tuning(self, rows):
        """rows: iterable of (class, bra_bucket, ket_bucket, variant, G[, regime]);
        regime 0 = small class launches, 1 = large, -1 (default) = both."""
        rows = [tuple(r) + ((-1,) if len(r) == 5 else ()) for r in rows]
        arr = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1, 6))
        self.check(self.L.mi_ctx_set_tuning(self.h, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            arr.shape[0]))

    def set_streams(self, n: int):
        self.check(self.L.mi_ctx_set_streams(self.h, int(n)))

    def set_fusion(self, on: bool):
        """Fuse the small class launches of plans built afterwards (default on)."""
        self.check(self.L.mi_ctx_set_fusion(self.h, 1 if on else 0))

    def plan(self, fb: FlatBasis, order=_lib.MI_ORDER_CHRR, precision=_lib.MI_PREC_FP64, tau=0.0,
             schwarz_cut=0.0, method=_lib.MI_METHOD_HGP) -> "Plan":
        return Plan(self, [fb], order, precision, tau, schwarz_cut, method)

    def batch_plan(self, fbs, order=_lib.MI_ORDER_CHRR, precision=_lib.MI_PREC_FP64, tau=0.0,
                   schwarz_cut=0.0, method=_lib.MI_METHOD_HGP) -> "Plan":
        """One plan over many molecules (outputs back to back, see Plan.mol_offsets)."""
        return Plan(self, list(fbs), order, precision, tau, schwarz_cut, method)


class Plan:
    """``mi_plan_create``: basis topology resident on the device; ``execute``
    writes all unique ERIs into a caller-owned device buffer (e.g. a torch
    tensor's ``data_ptr()``) on a given HIP stream."""

    def __init__(self, ctx: Context, fbs, order, precision, tau, schwarz_cut, method=_lib.MI_METHOD_HGP):
        self.ctx = ctx
        self.fbs = fbs
        self.fb = fbs[0]
        self._args = [_lib.BasisArgs(fb) for fb in fbs]
        arr = (_lib.MiBasis * len(fbs))(*[a.struct for a in self._args])
        o = _lib.opts(order, precision, tau, schwarz_cut, method=method)
        h = ctypes.c_void_p()
        if len(fbs) == 1:
            ctx.check(ctx.L.mi_plan_create(ctx.h, arr, ctypes.byref(o), ctypes.byref(h)))
        else:
            ctx.check(ctx.L.mi_batch_plan_create(ctx.h, arr, len(fbs), ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self.n_unique = int(ctx.L.mi_plan_out_len(h))

    def jk_execute(self, D_ptr: int, J_ptr: int, K_ptr: int, stream: int = 0):
        """Device-buffer Coulomb/exchange for a one-molecule plan."""
        self.ctx.check(self.ctx.L.mi_plan_jk_execute(self.h, ctypes.c_void_p(D_ptr), ctypes.c_void_p(J_ptr),
                                                     ctypes.c_void_p(K_ptr), ctypes.c_void_p(stream)))

    def pair_bounds(self):
        """(Q, V, mol) per shell pair after an execute with a Schwarz pass."""
        n = self.stats()["shell_pairs"]
        Q = np.empty(n)
        V = np.empty(n)
        mol = np.empty(n, dtype=np.int32)
        dp = ctypes.POINTER(ctypes.c_double)
        self.ctx.check(self.ctx.L.mi_plan_pair_bounds(self.h, Q.ctypes.data_as(dp), V.ctypes.data_as(dp),
                                                      mol.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return Q, V, mol

    def mol_offsets(self) -> np.ndarray:
        off = np.zeros(len(self.fbs) + 1, dtype=np.int64)
        self.ctx.check(self.ctx.L.mi_plan_mol_offsets(self.h, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return off

    def capture(self, out_ptr: int):
        """Capture one execute into a HIP graph (replay with ``replay``)."""
        self.ctx.check(self.ctx.L.mi_plan_graph_capture(self.h, ctypes.c_void_p(out_ptr), self.n_unique))

    def replay(self, stream: int = 0):
        self.ctx.check(self.ctx.L.mi_plan_graph_launch(self.h, ctypes.c_void_p(stream)))

    def execute(self, out_ptr: int, stream: int = 0):
        self.ctx.check(self.ctx.L.mi_plan_execute(self.h, ctypes.c_void_p(out_ptr), self.n_unique,
                                                  ctypes.c_void_p(stream)))

    def stats(self) -> dict:
        st = (ctypes.c_int64 * 7)()
        self.ctx.check(self.ctx.L.mi_plan_stats(self.h, st, 7))
        keys = ["nbf", "shell_pairs", "primitive_pairs", "shell_quartets", "primitive_quartets", "unique_eris",
                "launches"]
        return dict(zip(keys, [int(x) for x in st]))

    def launches(self) -> list:
        """Per class launch: ((la,lb,lc,ld), shell quartets, primitive quartets,
        blocks, kernel), kernel = "R" (generated register kernel), "W<nw>" (LDS
        wave kernel, nw wavefronts per quartet), "G<nw>" (global-table kernel,
        nw wavefronts per block), "T<nw>" (wave kernel, global op tables) or
        "X4" (double-double table kernel), "Y<nw>" (Rys kernel, nw wavefronts
        per block) or "Z<lg>" (Rys lane-group kernel, lg lanes per quartet)."""
        out = []
        n = self.ctx.L.mi_plan_nlaunch(self.h)
        for i in range(n):
            c4 = (ctypes.c_int32 * 4)()
            n3 = (ctypes.c_int64 * 3)()
            k2 = (ctypes.c_int32 * 2)()
            self.ctx.check(self.ctx.L.mi_plan_launch_info(self.h, i, c4, n3, k2))
            kern = "R" if k2[0] == 1 else {0: "W", 2: "G", 3: "T", 4: "X", 5: "Y", 6: "Z"}[k2[0]] + str(k2[1])
            out.append((tuple(c4), int(n3[0]), int(n3[1]), int(n3[2]), kern))
        return out

    def units(self) -> list:
        """Launch units of an execute: (kind, wavefronts per block, [launch
        indices]); kind 0 = one class launch, 1 = fused register kernels,
        2 = fused wave kernels, 3 = fused global-table kernels."""
        out = []
        for u in range(self.ctx.L.mi_plan_nunit(self.h)):
            k3 = (ctypes.c_int32 * 3)()
            mem = (ctypes.c_int32 * 64)()
            self.ctx.check(self.ctx.L.mi_plan_unit_info(self.h, u, k3, mem, 64))
            out.append((int(k3[0]), int(k3[1]), [int(mem[i]) for i in range(min(int(k3[2]), 64))]))
        return out

    def profile_units(self, out_ptr: int, stream: int = 0) -> list:
        """Per-unit durations (ms) of one execute as issued: [pair kernel, unit 0, ...]."""
        n = self.ctx.L.mi_plan_nunit(self.h) + 1
        ms = (ctypes.c_float * n)()
        self.ctx.check(self.ctx.L.mi_plan_profile_units(self.h, ctypes.c_void_p(out_ptr), self.n_unique,
                                                        ctypes.c_void_p(stream), ms))
        return [float(x) for x in ms]

    def profile(self, out_ptr: int, stream: int = 0) -> list:
        """Per-launch durations (ms) of one execute: [pair kernel, class 0, ...]."""
        n = self.ctx.L.mi_plan_nlaunch(self.h) + 1
        ms = (ctypes.c_float * n)()
        self.ctx.check(self.ctx.L.mi_plan_profile(self.h, ctypes.c_void_p(out_ptr), self.n_unique,
                                                  ctypes.c_void_p(stream), ms))
        return [float(x) for x in ms]

    def balance(self, out_ptr: int, stream: int = 0):
        """One profiled execute, then stream lanes reassigned by measured
        launch times (mi_plan_balance)."""
        self.ctx.check(self.ctx.L.mi_plan_balance(self.h, ctypes.c_void_p(out_ptr), self.n_unique,
                                                  ctypes.c_void_p(stream)))

    def jk_profile(self, D_ptr: int, J_ptr: int, K_ptr: int, stream: int = 0) -> list:
        """Per-launch durations (ms) of one Coulomb/exchange execute."""
        n = self.ctx.L.mi_plan_nlaunch(self.h) + 1
        ms = (ctypes.c_float * n)()
        self.ctx.check(self.ctx.L.mi_plan_jk_profile(self.h, ctypes.c_void_p(D_ptr), ctypes.c_void_p(J_ptr),
                                                     ctypes.c_void_p(K_ptr), ctypes.c_void_p(stream), ms))
        return [float(x) for x in ms]

    def close(self):
        if getattr(self, "h", None):
            self.ctx.L.mi_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def default_context() -> Context:
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx


def _flat(bfs) -> FlatBasis:
    return bfs if isinstance(bfs, FlatBasis) else flatten(bfs)


def all_twoe_ints_chrr(bfs) -> np.ndarray:
    """GPU ``all_twoe_ints_chrr`` (src/HGP.jl:22-34), 0-based slots."""
    return default_context().eri_unique(_flat(bfs), order=_lib.MI_ORDER_CHRR)


def all_twoe_ints(bfs) -> np.ndarray:
    """GPU ``all_twoe_ints`` (src/ERI.jl:104-112), ``iindex`` slots (0-based)."""
    return default_context().eri_unique(_flat(bfs), order=_lib.MI_ORDER_IINDEX)
