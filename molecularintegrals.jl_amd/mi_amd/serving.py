"""Serving loop over the engine's continuous-batching server (mi_server_* in
include/mi_hgp.h).

``Server`` wraps the C engine.  ``serve`` replays a request schedule (arrival
times from a seeded Poisson process, ``configs.serving_requests``) in real
time: requests are submitted when their arrival time has passed, the engine
is stepped continuously, and each request's latency is its completion time
minus its arrival time.  Inputs (densities) and output buffers are staged on
the device before the clock starts, as a client holding them in HBM would.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np

from . import _lib
from .configs import Request


class Server:
    """One engine on one device; single host thread."""

    def __init__(self, device: int = 0, max_inflight: int = 4, max_batch: int = 64, hbm_budget: int = 64 << 30):
        self.L = _lib.load()
        h = ctypes.c_void_p()
        o = _lib.MiServerOpts(max_inflight, max_batch, hbm_budget)
        rc = self.L.mi_server_create(int(device), ctypes.byref(o), ctypes.byref(h))
        if rc != 0:
            raise _lib.MiError(rc, f"mi_server_create(device={device}) failed")
        self.h = h

    def check(self, rc):
        if rc != 0:
            raise _lib.MiError(rc, self.L.mi_server_last_error(self.h).decode())

    def submit(self, fb, kind: int, out_ptr: int = 0, D_ptr: int = 0, J_ptr: int = 0, K_ptr: int = 0,
               order=_lib.MI_ORDER_CHRR, precision=_lib.MI_PREC_FP64, tau=0.0, schwarz_cut=0.0,
               method=_lib.MI_METHOD_HGP) -> int:
        args = _lib.BasisArgs(fb)
        o = _lib.opts(order, precision, tau, schwarz_cut, method=method)
        rid = ctypes.c_int64()
        vp = ctypes.c_void_p
        self.check(self.L.mi_server_submit(self.h, ctypes.byref(args.struct), ctypes.byref(o), int(kind),
                                           vp(D_ptr or None), vp(out_ptr or None), vp(J_ptr or None),
                                           vp(K_ptr or None), ctypes.byref(rid)))
        return int(rid.value)

    def step(self):
        self.check(self.L.mi_server_step(self.h))

    def poll(self, rid: int) -> int:
        st = ctypes.c_int32()
        self.check(self.L.mi_server_poll(self.h, int(rid), ctypes.byref(st)))
        return int(st.value)

    def release(self, rid: int):
        self.check(self.L.mi_server_release(self.h, int(rid)))

    def drain(self):
        self.check(self.L.mi_server_drain(self.h))

    def stats(self) -> dict:
        st = (ctypes.c_int64 * 8)()
        self.check(self.L.mi_server_stats(self.h, st, 8))
        keys = ["queued", "running", "done", "failed", "batches", "bytes_in_use", "peak_bytes", "inflight_batches"]
        return dict(zip(keys, [int(x) for x in st]))

    def close(self):
        if getattr(self, "h", None):
            self.L.mi_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ServeResult:
    latency_s: Dict[int, float] = field(default_factory=dict)   # request rid -> latency
    makespan_s: float = 0.0
    unique_eris: int = 0
    outputs: Dict[int, tuple] = field(default_factory=dict)     # rid -> device tensors
    stats: dict = field(default_factory=dict)
    timeline: List[tuple] = field(default_factory=list)         # (t, batches in flight, requests running)

    def overlap(self) -> dict:
        """Batches in flight over the run: time-weighted mean, maximum, and the
        fraction of the makespan with two or more batches on the device."""
        if len(self.timeline) < 2:
            return {"mean_batches_in_flight": 0.0, "max_batches_in_flight": 0, "frac_time_2plus_batches": 0.0}
        t = np.array([x[0] for x in self.timeline])
        b = np.array([x[1] for x in self.timeline], dtype=float)
        dt = np.diff(t)
        tot = float(dt.sum()) or 1.0
        return {"mean_batches_in_flight": float((b[:-1] * dt).sum() / tot),
                "max_batches_in_flight": int(b.max()),
                "frac_time_2plus_batches": float(dt[b[:-1] >= 2].sum() / tot)}

    def p(self, q: float) -> float:
        return float(np.quantile(np.array(list(self.latency_s.values())), q))

    @property
    def throughput(self) -> float:
        return self.unique_eris / self.makespan_s if self.makespan_s > 0 else 0.0


def mixed_options(kind: str) -> dict:
    """The mixed form: FP32 VRR where the VRR-output bound product is below
    1e-8, quartets with Schwarz bound product below 1e-12 skipped."""
    return dict(precision=_lib.MI_PREC_MIXED, tau=1e-8, schwarz_cut=1e-12)


def serve(requests: List[Request], mixed: bool = False, max_inflight: int = 4, max_batch: int = 64,
          hbm_budget: int = 64 << 30, time_scale: float = 1.0, device: int = 0) -> ServeResult:
    """Replay ``requests`` (sorted by arrival) through a fresh server.
    time_scale stretches the arrival times (2.0 = half the rate)."""
    import torch

    dev = torch.device("cuda", device)
    srv = Server(device, max_inflight, max_batch, hbm_budget)
    res = ServeResult()
    staged = []
    for r in requests:
        fb = r.workload.flat()
        n = fb.nbf
        res.unique_eris += n * (n + 1) * (n * n + n + 2) // 8
        if r.kind == "eri":
            out = torch.empty(n * (n + 1) * (n * n + n + 2) // 8, dtype=torch.float64, device=dev)
            staged.append((r, fb, (out,)))
        else:
            D = torch.from_numpy(np.ascontiguousarray(r.density)).to(dev)
            J = torch.empty(n, n, dtype=torch.float64, device=dev)
            K = torch.empty_like(J)
            staged.append((r, fb, (D, J, K)))
    torch.cuda.synchronize(dev)
    pending = list(staged)
    running = {}
    t0 = time.perf_counter()
    while pending or running:
        now = time.perf_counter() - t0
        while pending and pending[0][0].arrival_s * time_scale <= now:
            r, fb, bufs = pending.pop(0)
            kw = mixed_options(r.kind) if mixed else {}
            if r.kind == "eri":
                rid = srv.submit(fb, _lib.MI_REQ_ERI, out_ptr=bufs[0].data_ptr(), **kw)
            else:
                rid = srv.submit(fb, _lib.MI_REQ_JK, D_ptr=bufs[0].data_ptr(), J_ptr=bufs[1].data_ptr(),
                                 K_ptr=bufs[2].data_ptr(), **kw)
            running[rid] = r
            res.outputs[r.rid] = bufs
        srv.step()
        now = time.perf_counter() - t0
        st = srv.stats()
        if not res.timeline or res.timeline[-1][1:] != (st["inflight_batches"], st["running"]):
            res.timeline.append((now, st["inflight_batches"], st["running"]))
        for rid in list(running):
            st = srv.poll(rid)
            if st == _lib.MI_REQ_DONE:
                r = running.pop(rid)
                res.latency_s[r.rid] = now - r.arrival_s * time_scale
                srv.release(rid)
            elif st == _lib.MI_REQ_FAILED:
                raise _lib.MiError(_lib.MI_EDEVICE, srv.L.mi_server_last_error(srv.h).decode())
        if not running and pending:
            wait = pending[0][0].arrival_s * time_scale - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(min(wait, 1e-3))
    res.makespan_s = time.perf_counter() - t0
    res.timeline.append((res.makespan_s, 0, 0))
    res.stats = srv.stats()
    srv.close()
    return res
