"""ctypes binding of libmi_hgp.so (declarations: include/mi_hgp.h).

The library is the only compute path: if it is missing or fails to load, every
entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("MI_HGP_LIB", os.path.join(PKG_ROOT, "libmi_hgp.so"))

MI_OK, MI_EINVAL, MI_EUNSUPPORTED, MI_EDEVICE, MI_ENOMEM, MI_ESIZE = range(6)
MI_ORDER_IINDEX, MI_ORDER_CHRR, MI_ORDER_JK = 0, 1, 2  # MI_ORDER_JK: plan for J/K builds only
MI_PREC_FP64, MI_PREC_MIXED = 0, 1
MI_METHOD_HGP, MI_METHOD_RYS = 0, 1

# Every symbol include/mi_hgp.h declares (checked by the CPU test suite).
EXPORTED = ["mi_create", "mi_destroy", "mi_last_error", "mi_lmax", "mi_n_unique", "mi_eri_unique",
            "mi_rys_rule", "mi_rys_nmax",
            "mi_eri_unique_noalloc",
            "mi_plan_create", "mi_plan_execute", "mi_plan_destroy", "mi_plan_stats", "mi_plan_nlaunch",
            "mi_plan_launch_info", "mi_plan_profile", "mi_plan_jk_profile", "mi_plan_balance", "mi_ctx_set_tuning", "mi_ctx_set_streams",
            "mi_ctx_set_fusion", "mi_plan_nunit", "mi_plan_unit_info", "mi_plan_profile_units",
            "mi_batch_plan_create", "mi_plan_out_len", "mi_plan_mol_offsets", "mi_plan_graph_capture",
            "mi_plan_graph_launch", "mi_jk", "mi_plan_jk_execute", "mi_plan_pair_bounds",
            "mi_server_create", "mi_server_destroy", "mi_server_last_error", "mi_server_submit", "mi_server_step",
            "mi_server_poll", "mi_server_release", "mi_server_drain", "mi_server_stats"]

MI_REQ_ERI, MI_REQ_JK = 0, 1
MI_REQ_QUEUED, MI_REQ_RUNNING, MI_REQ_DONE, MI_REQ_FAILED = 0, 1, 2, 3


class MiBasis(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int32),
                ("shell_xyz", ctypes.POINTER(ctypes.c_double)),
                ("shell_l", ctypes.POINTER(ctypes.c_int32)),
                ("shell_prim_off", ctypes.POINTER(ctypes.c_int32)),
                ("prim_exp", ctypes.POINTER(ctypes.c_double)),
                ("prim_coef", ctypes.POINTER(ctypes.c_double)),
                ("prim_norm", ctypes.POINTER(ctypes.c_double)),
                ("nbf", ctypes.c_int32),
                ("bf_scale", ctypes.POINTER(ctypes.c_double))]


class MiOpts(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32),
                ("precision", ctypes.c_int32),
                ("tau", ctypes.c_double),
                ("schwarz_cut", ctypes.c_double),
                ("stream", ctypes.c_void_p),
                ("method", ctypes.c_int32)]


class MiServerOpts(ctypes.Structure):
    _fields_ = [("max_inflight", ctypes.c_int32),
                ("max_batch", ctypes.c_int32),
                ("hbm_budget", ctypes.c_int64)]


class MiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libmi_hgp error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libmi_hgp.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64.so.7 (same
    # soname as /opt/rocm's).  Whichever loads first serves both, and torch
    # cannot run on a newer runtime than it was built for, so let torch load
    # its runtime first when it is installed.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libmi_hgp.so not found at {LIB_PATH}; build it with "
                           f"python molecularintegrals.jl_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.mi_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.mi_create.restype = ctypes.c_int
    L.mi_destroy.argtypes = [vp]
    L.mi_destroy.restype = None
    L.mi_last_error.argtypes = [vp]
    L.mi_last_error.restype = ctypes.c_char_p
    L.mi_lmax.argtypes = []
    L.mi_lmax.restype = ctypes.c_int32
    L.mi_n_unique.argtypes = [ctypes.c_int32]
    L.mi_n_unique.restype = ctypes.c_int64
    L.mi_rys_rule.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double)]
    L.mi_rys_rule.restype = ctypes.c_int
    L.mi_rys_nmax.argtypes = []
    L.mi_rys_nmax.restype = ctypes.c_int32
    L.mi_eri_unique.argtypes = [vp, ctypes.POINTER(MiBasis), ctypes.POINTER(MiOpts),
                                ctypes.POINTER(ctypes.c_double), ctypes.c_int64]
    L.mi_eri_unique.restype = ctypes.c_int
    L.mi_eri_unique_noalloc.argtypes = [vp, ctypes.POINTER(MiBasis), ctypes.POINTER(MiOpts),
                                vp, ctypes.c_int64]
    L.mi_eri_unique_noalloc.restype = ctypes.c_int
    L.mi_plan_create.argtypes = [vp, ctypes.POINTER(MiBasis), ctypes.POINTER(MiOpts), ctypes.POINTER(vp)]
    L.mi_plan_create.restype = ctypes.c_int
    L.mi_plan_execute.argtypes = [vp, vp, ctypes.c_int64, vp]
    L.mi_plan_execute.restype = ctypes.c_int
    L.mi_plan_destroy.argtypes = [vp]
    L.mi_plan_destroy.restype = None
    L.mi_plan_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    L.mi_plan_stats.restype = ctypes.c_int
    L.mi_plan_nlaunch.argtypes = [vp]
    L.mi_plan_nlaunch.restype = ctypes.c_int
    L.mi_plan_launch_info.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]
    L.mi_plan_launch_info.restype = ctypes.c_int
    L.mi_plan_profile.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.POINTER(ctypes.c_float)]
    L.mi_plan_jk_profile.argtypes = [vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_float)]
    L.mi_plan_jk_profile.restype = ctypes.c_int
    L.mi_plan_balance.argtypes = [vp, vp, ctypes.c_int64, vp]
    L.mi_plan_balance.restype = ctypes.c_int
    L.mi_plan_profile.restype = ctypes.c_int
    L.mi_ctx_set_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    L.mi_ctx_set_tuning.restype = ctypes.c_int
    L.mi_ctx_set_streams.argtypes = [vp, ctypes.c_int32]
    L.mi_ctx_set_streams.restype = ctypes.c_int
    L.mi_ctx_set_fusion.argtypes = [vp, ctypes.c_int32]
    L.mi_ctx_set_fusion.restype = ctypes.c_int
    L.mi_plan_nunit.argtypes = [vp]
    L.mi_plan_nunit.restype = ctypes.c_int
    L.mi_plan_unit_info.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    L.mi_plan_unit_info.restype = ctypes.c_int
    L.mi_plan_profile_units.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.POINTER(ctypes.c_float)]
    L.mi_plan_profile_units.restype = ctypes.c_int
    L.mi_batch_plan_create.argtypes = [vp, ctypes.POINTER(MiBasis), ctypes.c_int32, ctypes.POINTER(MiOpts),
                                       ctypes.POINTER(vp)]
    L.mi_batch_plan_create.restype = ctypes.c_int
    L.mi_plan_out_len.argtypes = [vp]
    L.mi_plan_out_len.restype = ctypes.c_int64
    L.mi_plan_mol_offsets.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.mi_plan_mol_offsets.restype = ctypes.c_int
    L.mi_plan_graph_capture.argtypes = [vp, vp, ctypes.c_int64]
    L.mi_plan_graph_capture.restype = ctypes.c_int
    L.mi_plan_graph_launch.argtypes = [vp, vp]
    L.mi_plan_graph_launch.restype = ctypes.c_int
    dpp = ctypes.POINTER(ctypes.c_double)
    L.mi_jk.argtypes = [vp, ctypes.POINTER(MiBasis), ctypes.POINTER(MiOpts), dpp, dpp, dpp]
    L.mi_jk.restype = ctypes.c_int
    L.mi_plan_jk_execute.argtypes = [vp, vp, vp, vp, vp]
    L.mi_plan_jk_execute.restype = ctypes.c_int
    L.mi_plan_pair_bounds.argtypes = [vp, dpp, dpp, ctypes.POINTER(ctypes.c_int32)]
    L.mi_plan_pair_bounds.restype = ctypes.c_int
    L.mi_server_create.argtypes = [ctypes.c_int32, ctypes.POINTER(MiServerOpts), ctypes.POINTER(vp)]
    L.mi_server_create.restype = ctypes.c_int
    L.mi_server_destroy.argtypes = [vp]
    L.mi_server_destroy.restype = None
    L.mi_server_last_error.argtypes = [vp]
    L.mi_server_last_error.restype = ctypes.c_char_p
    L.mi_server_submit.argtypes = [vp, ctypes.POINTER(MiBasis), ctypes.POINTER(MiOpts), ctypes.c_int32, vp, vp, vp,
                                   vp, ctypes.POINTER(ctypes.c_int64)]
    L.mi_server_submit.restype = ctypes.c_int
    L.mi_server_step.argtypes = [vp]
    L.mi_server_step.restype = ctypes.c_int
    L.mi_server_poll.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
    L.mi_server_poll.restype = ctypes.c_int
    L.mi_server_release.argtypes = [vp, ctypes.c_int64]
    L.mi_server_release.restype = ctypes.c_int
    L.mi_server_drain.argtypes = [vp]
    L.mi_server_drain.restype = ctypes.c_int
    L.mi_server_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]
    L.mi_server_stats.restype = ctypes.c_int
    _lib = L
    return L


class BasisArgs:
    """Keeps the numpy arrays alive while the C struct points into them."""

    def __init__(self, fb):
        arrs = fb.arrays()
        self._keep = arrs
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        self.struct = MiBasis(fb.nshell, dp(arrs[0]), ip(arrs[1]), ip(arrs[2]), dp(arrs[3]), dp(arrs[4]),
                              dp(arrs[5]), fb.nbf, dp(arrs[6]))


def opts(order=MI_ORDER_CHRR, precision=MI_PREC_FP64, tau=0.0, schwarz_cut=0.0, stream=None,
         method=MI_METHOD_HGP):
    return MiOpts(order, precision, tau, schwarz_cut, stream, method)


def rys_rule(n: int, T: float):
    """The n-point Rys rule at T as the Rys kernel evaluates it (host copy of
    its tables): roots x = t^2 and weights (mi_rys_rule)."""
    L = load()
    x = np.empty(n)
    w = np.empty(n)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.mi_rys_rule(int(n), float(T), x.ctypes.data_as(dp), w.ctypes.data_as(dp))
    if rc != MI_OK:
        raise MiError(rc, f"mi_rys_rule(n={n}, T={T})")
    return x, w
