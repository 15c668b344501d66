"""Synthetic workload generators on the host (numpy), for the benchmark and
the tests.

``general_sparse``: the general sparse workload -- a nonsymmetric CSR matrix
built as a random band of a given width (``per_row`` distinct off-diagonal
offsets per row drawn uniformly from [-width/2, width/2] minus 0, plus the
diagonal), values lognormal(0, 1) with random signs, a fraction of the
diagonal entries scaled by 1e-3 (so DB matching is needed), then a
symmetric random permutation.  At N=500,000, width 300 and 11 offsets per
row it has ~11.85 entries per row (entries whose column falls outside
[0, N) are dropped).

The banded workloads (off-diagonals U[-1,1), diagonal = dominance x row
off-diagonal abs-sum) are generated on the GPU by the library itself
(``BandedMatrix.generate``).
"""
import numpy as np


def general_sparse(n, width=300, per_row=11, seed=20, weak_seed=21, weak_diag=0.2, chunk=1 << 15):
    """(row_offsets int32 (n+1), column_indices int32 (nnz), values float64
    (nnz)) of the general sparse workload, sorted by (row, column).

    seed draws the pattern, the values, their signs and the permutation;
    weak_seed draws which diagonal entries are scaled by 1e-3 (each with
    probability weak_diag)."""
    rng = np.random.default_rng(seed)
    h = width // 2
    cand = np.r_[-h:0, 1:h + 1]
    m = min(per_row, cand.size)
    offs = np.empty((n, m), dtype=np.int64)
    for r0 in range(0, n, chunk):
        r1 = min(n, r0 + chunk)
        # m distinct offsets per row: the m smallest of i.i.d. uniform keys
        keys = rng.random((r1 - r0, cand.size))
        offs[r0:r1] = cand[np.argpartition(keys, m - 1, axis=1)[:, :m]]
    rows = np.repeat(np.arange(n, dtype=np.int64), m + 1)
    cols = np.concatenate([np.arange(n, dtype=np.int64)[:, None] + offs,
                           np.arange(n, dtype=np.int64)[:, None]], axis=1).reshape(-1)
    keep = (cols >= 0) & (cols < n)
    rows, cols = rows[keep], cols[keep]
    vals = rng.lognormal(0.0, 1.0, rows.size) * np.where(rng.random(rows.size) < 0.5, -1.0, 1.0)
    weak = np.random.default_rng(weak_seed).random(n) < weak_diag
    diag = rows == cols
    vals[diag & weak[rows]] *= 1e-3
    p = rng.permutation(n)
    rows, cols = p[rows], p[cols]
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rp[1:])
    return rp.astype(np.int32), cols.astype(np.int32), vals


def uniform_vector(n, seed):
    """x_true ~ U[-1, 1) from a numpy generator (host)."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, n)
