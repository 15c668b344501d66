"""Writes the committed golden fixtures under tests/golden/.

The fixtures pin this project's synthetic workload generator (xoshiro256**
per row, seeded by splitmix64; the definition lives in oracle/sap_oracle.cpp
and in the HIP generator kernel) so that the CPU oracle and the GPU generator
can be compared bit for bit, and a change of definition is caught.  The
reference repository holds no golden vectors or fixtures of its own.

Run:  make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402


def main():
    band = O.gen_banded(40, 3, 1.2, 0)
    x = O.gen_vector(40, 1)
    b = O.band_spmv(band, x)
    band2 = O.gen_banded(300, 12, 0.8, 10)
    np.savez(os.path.join(HERE, "generator_small.npz"), band_n40_k3_d12_s0=band, vec_n40_s1=x,
             spmv_n40=b, band_n300_k12_d08_s10=band2)
    print("wrote generator_small.npz")


if __name__ == "__main__":
    main()
