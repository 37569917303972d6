"""The library's host code (plan builder, launch units, graphs, J/K, serving
engine, allocation-free entry point, error paths) built with
AddressSanitizer and UndefinedBehaviorSanitizer on the host side only
(tests/asan/Makefile; device code as in the library) and driven through the
C ABI by tests/asan/mi_asan_driver.cpp on 20 molecules of the batched recipe
and a 4-atom cluster-recipe molecule with f shells.  The driver checks its
results against the one-shot path at the parity bar; any sanitizer report
ends it with a nonzero status."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "asan", "mi_asan_driver")
TUNING = os.path.join(os.path.dirname(HERE), "molecularintegrals.jl_amd", "hgp_tuning.txt")


def _write_bases(path, fbs):
    with open(path, "wb") as f:
        f.write(np.int32(len(fbs)).tobytes())
        for fb in fbs:
            f.write(np.array([fb.nshell, fb.prim_exp.size, fb.nbf], dtype=np.int32).tobytes())
            xyz, sl, off, ex, co, pn, sc = fb.arrays()
            for a in (xyz, sl, off, ex, co, pn, sc):
                f.write(np.ascontiguousarray(a).tobytes())


def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("host-sanitized driver not built (make -C tests/asan)")
    from mi_amd import configs
    fbs = [m.flat() for m in configs.batch_molecules(3, nmol=20)]
    rng = np.random.Generator(np.random.PCG64(21))
    fbs.append(configs.random_molecule(rng, 4, configs.CLUSTER_HEAVY, configs.CLUSTER_LIGHT).flat())
    path = tmp_path / "bases.bin"
    _write_bases(path, fbs)
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:exitcode=23",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=24",
               MI_HGP_TUNING=TUNING)
    print(f"running {DRIVER} on {len(fbs)} molecules", flush=True)
    r = subprocess.run([DRIVER, str(path)], env=env, capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr[-6000:]
    assert "mi_asan_driver: ok" in r.stdout
