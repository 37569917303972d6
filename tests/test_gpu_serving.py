"""Serving engine (continuous batching) on the GPU, through the C ABI.

The bar: every request's result is the same whether it ran alone (one-shot
mi_eri_unique / mi_jk) or merged and overlapped with others in the engine --
unique ERIs within 1e-12 abs / 1e-10 rel (FP64: bit for bit, the per-quartet
arithmetic does not depend on the batch), J/K within 1e-11 relative (FP64
atomics).  The mixed form stays within 1e-10 Eh of the FP64 form.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_requests(seed, n_eri=5, n_jk=3):
    from mi_amd import configs
    from mi_amd.configs import Request
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs = []
    for k in range(n_eri):
        wl = configs.random_molecule(rng, int(rng.integers(2, 6)), configs.SPLIT_HEAVY, configs.SPLIT_LIGHT)
        reqs.append(Request(len(reqs), 0.0, wl, "eri"))
    for k in range(n_jk):
        wl = configs.random_molecule(rng, int(rng.integers(2, 5)), configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT)
        n = wl.nbf
        d = rng.normal(size=(n, n)) / n
        reqs.append(Request(len(reqs), 0.0, wl, "jk", 0.5 * (d + d.T)))
    order = rng.permutation(len(reqs))
    return [reqs[i] for i in order]


def _alone(ctx, r, **kw):
    fb = r.workload.flat()
    if r.kind == "eri":
        return (ctx.eri_unique(fb, **kw),)
    return ctx.jk(fb, r.density, **kw)


@pytest.mark.parametrize("max_inflight,max_batch", [(4, 64), (2, 2), (1, 1)])
def test_results_alone_equal_in_the_mix(ctx, max_inflight, max_batch):
    """All requests submitted at once: ERI requests merge into batched plans,
    batches overlap on separate streams."""
    import torch
    from mi_amd import _lib
    from mi_amd.serving import Server
    reqs = _small_requests(7)
    srv = Server(0, max_inflight=max_inflight, max_batch=max_batch)
    bufs, ids = {}, {}
    for r in reqs:
        fb = r.workload.flat()
        n = fb.nbf
        if r.kind == "eri":
            out = torch.full((n * (n + 1) * (n * n + n + 2) // 8,), float("nan"), dtype=torch.float64, device="cuda")
            bufs[r.rid] = (out,)
            ids[r.rid] = srv.submit(fb, _lib.MI_REQ_ERI, out_ptr=out.data_ptr())
        else:
            D = torch.from_numpy(r.density).cuda()
            J = torch.empty(n, n, dtype=torch.float64, device="cuda")
            K = torch.empty_like(J)
            bufs[r.rid] = (D, J, K)
            ids[r.rid] = srv.submit(fb, _lib.MI_REQ_JK, D_ptr=D.data_ptr(), J_ptr=J.data_ptr(), K_ptr=K.data_ptr())
    srv.drain()
    st = srv.stats()
    print(f"inflight={max_inflight} batch={max_batch}: {st}")
    assert st["done"] == len(reqs) and st["failed"] == 0
    n_eri = sum(r.kind == "eri" for r in reqs)
    if max_batch > 1:
        assert st["batches"] < len(reqs)  # ERI requests were merged
    else:
        assert st["batches"] == len(reqs)
    for r in reqs:
        assert srv.poll(ids[r.rid]) == _lib.MI_REQ_DONE
        ref = _alone(ctx, r)
        if r.kind == "eri":
            got = bufs[r.rid][0].cpu().numpy()
            d = np.abs(got - ref[0])
            assert np.all((d <= 1e-12) | (d <= 1e-10 * np.abs(ref[0]))), f"request {r.rid}: max |d| {d.max():.2e}"
            print(f"request {r.rid} ERI nbf={r.workload.nbf}: bitwise equal to alone: {np.array_equal(got, ref[0])}")
        else:
            J, K = bufs[r.rid][1].cpu().numpy(), bufs[r.rid][2].cpu().numpy()
            scale = max(1.0, np.abs(ref[0]).max())
            assert np.abs(J - ref[0]).max() <= 1e-11 * scale and np.abs(K - ref[1]).max() <= 1e-11 * scale
    assert n_eri > 0
    srv.close()


def test_admission_by_device_memory():
    """Requests are admitted by the device bytes of their plans (the outputs
    are the callers' buffers).  With a budget of 1.5 x the largest small
    request, the small requests all complete without the engine ever holding
    more than the budget, and a request beyond it fails cleanly without
    blocking the queue."""
    import torch
    from mi_amd import _lib, configs
    from mi_amd.serving import Server
    rng = np.random.Generator(np.random.PCG64(3))
    wls = [configs.random_molecule(rng, 3, configs.SPLIT_HEAVY, configs.SPLIT_LIGHT) for _ in range(4)]
    big = configs.random_molecule(rng, 14, configs.SPLIT_HEAVY, configs.SPLIT_LIGHT)
    outs = []
    for w in wls + [big]:
        n = w.nbf
        outs.append(torch.empty(n * (n + 1) * (n * n + n + 2) // 8, dtype=torch.float64, device="cuda"))
    # the largest small request's bytes: the peak of a server running them one at a time
    probe = Server(0, max_inflight=1, max_batch=1)
    for w, o in zip(wls, outs):
        probe.submit(w.flat(), _lib.MI_REQ_ERI, out_ptr=o.data_ptr())
    probe.drain()
    budget = int(1.5 * probe.stats()["peak_bytes"])
    probe.close()
    srv = Server(0, max_inflight=4, max_batch=8, hbm_budget=budget)
    ids = [srv.submit(w.flat(), _lib.MI_REQ_ERI, out_ptr=o.data_ptr()) for w, o in zip(wls + [big], outs)]
    with pytest.raises(_lib.MiError):
        srv.drain()  # the big request fails: MI_EDEVICE from drain
    st = srv.stats()
    print(f"budget {budget} B: {st}")
    assert st["done"] == len(wls) and st["failed"] == 1
    assert st["peak_bytes"] <= budget
    assert srv.poll(ids[-1]) == _lib.MI_REQ_FAILED
    from mi_amd import Context
    c = Context(0)
    for w, o in zip(wls, outs):
        assert np.array_equal(o.cpu().numpy(), c.eri_unique(w.flat()))
    c.close()
    srv.close()


def test_poisson_schedule_latency_and_mixed_error():
    """A seeded Poisson schedule of ERI and J/K requests replayed in real
    time: every request completes, latencies are reported, and the mixed form
    (FP32 VRR below 1e-8, screening at 1e-12) is within 1e-10 of FP64."""
    from mi_amd import configs
    from mi_amd.serving import serve
    reqs = configs.serving_requests(seed=5, nreq=12, rate_per_s=400.0, natoms_lo=2, natoms_hi=10)
    plain = serve(reqs, mixed=False)
    mixed = serve(reqs, mixed=True)
    for res, name in ((plain, "fp64"), (mixed, "mixed")):
        assert len(res.latency_s) == len(reqs)
        assert all(v > 0 for v in res.latency_s.values())
        print(f"{name}: p50 {res.p(0.5) * 1e3:.2f} ms p99 {res.p(0.99) * 1e3:.2f} ms "
              f"throughput {res.throughput:.3e} unique ERIs/s stats {res.stats}")
    err = 0.0
    for r in reqs:
        a = [t.cpu().numpy() for t in plain.outputs[r.rid]]
        b = [t.cpu().numpy() for t in mixed.outputs[r.rid]]
        k = 0 if r.kind == "eri" else 1
        err = max(err, max(np.abs(x - y).max() for x, y in zip(a[k:], b[k:])))
    print(f"mixed vs FP64 max |d| = {err:.2e}")
    assert err <= 1e-10


def test_step_does_not_wait_for_a_mixed_batch():
    """A mixed, screened J/K request (a 1,000-function cluster-recipe
    molecule, seconds of device work and a Schwarz bounds pass with a device
    read-back at plan build): every mi_server_step call returns at once while
    the request is sized, built and run by the engine's workers -- the
    request is seen running, and no step call lasts more than 50 ms."""
    import time
    import torch
    from mi_amd import _lib, configs
    from mi_amd.serving import Server
    rng = np.random.Generator(np.random.PCG64(12))
    wl = configs.random_molecule(rng, 44, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT)
    fb = wl.flat()
    n = fb.nbf
    D = torch.from_numpy(0.5 * (lambda d: d + d.T)(rng.normal(size=(n, n)) / n)).cuda()
    J = torch.empty(n, n, dtype=torch.float64, device="cuda")
    K = torch.empty_like(J)
    srv = Server(0)
    rid = srv.submit(fb, _lib.MI_REQ_JK, D_ptr=D.data_ptr(), J_ptr=J.data_ptr(), K_ptr=K.data_ptr(),
                     precision=_lib.MI_PREC_MIXED, tau=1e-8, schwarz_cut=1e-12)
    seen, worst, t_end = set(), 0.0, time.time() + 60
    while time.time() < t_end:
        t = time.perf_counter()
        srv.step()
        worst = max(worst, time.perf_counter() - t)
        st = srv.poll(rid)
        seen.add(st)
        if st in (_lib.MI_REQ_DONE, _lib.MI_REQ_FAILED):
            break
        time.sleep(1e-3)
    print(f"{n} functions: states seen {sorted(seen)}, longest step {worst * 1e3:.2f} ms")
    assert srv.poll(rid) == _lib.MI_REQ_DONE
    assert _lib.MI_REQ_RUNNING in seen
    assert worst < 0.05
    srv.close()
