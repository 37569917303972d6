"""Writes the golden fixtures of the HGP ERI path from the CPU oracle.

ethane_minimal_seed0.npz        full 9,316-slot array (chrr order) of the
                                16-function ethane-like workload, seed 0.
ethane_split_valence_seed0.npz  10,000 sampled slots + sum/abs-sum + SHA-256 of
                                the 1,675,365-slot array (chrr order), seed 0.
The oracle is pinned separately against the reference's known-answer tests
(tests/test_oracle.py); these fixtures freeze its output so that a change to
the oracle or to the workload generator is caught.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "molecularintegrals.jl_amd"))

from oracle import oracle as O  # noqa: E402
from mi_amd import configs  # noqa: E402


def main():
    fb0 = configs.ethane_minimal(0).flat()
    e0 = O.eri_unique_hgp(fb0, order=O.ORDER_CHRR)
    np.savez(os.path.join(HERE, "ethane_minimal_seed0.npz"), values=e0)
    fb1 = configs.ethane_split_valence(0).flat()
    e1 = O.eri_unique_hgp(fb1, order=O.ORDER_CHRR)
    rng = np.random.Generator(np.random.PCG64(12345))
    slots = np.sort(rng.choice(e1.size, size=10000, replace=False))
    np.savez(os.path.join(HERE, "ethane_split_valence_seed0.npz"), slots=slots, values=e1[slots],
             sum=e1.sum(), abs_sum=np.abs(e1).sum(), sha256=hashlib.sha256(e1.tobytes()).hexdigest())
    print("wrote fixtures", e0.size, e1.size)


if __name__ == "__main__":
    main()
