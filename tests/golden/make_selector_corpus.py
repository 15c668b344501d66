"""Writes tests/golden/selector_corpus.npz (the committed file holds matrices
0-199: two runs, 100 0 and 100 100, concatenated): features and measured
time-to-solution of every option set of the selector's grid for the
corpus matrices (saplibrary_amd.selector.generate_corpus), measured by the
solver on the GPU this script runs on.

    python tests/golden/make_selector_corpus.py [count] [out.npz] [start]

(several runs with different start indices can be concatenated).
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import saplibrary_amd as sap  # noqa: E402
from saplibrary_amd import selector as S  # noqa: E402


def main():
    count = int(sys.argv[1]) if len(sys.argv) > 1 else 240
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(ROOT, "tests", "golden", "selector_corpus.npz")
    start = int(sys.argv[3]) if len(sys.argv) > 3 else 0

    def save(rows, F, T, OK, kinds):
        # written after every matrix, so a run cut short keeps what it measured
        np.savez_compressed(out, features=F, times=T, ok=OK, kinds=kinds, grid=np.array(S.GRID),
                            default_index=S.DEFAULT_INDEX)

    S.generate_corpus(sap, torch, count, start=start, log=lambda m: print(m, flush=True), on_row=save)
    print(f"wrote {out}", flush=True)


if __name__ == "__main__":
    main()
