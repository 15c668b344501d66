"""Sparse test matrices shared by the CSR tests."""
import numpy as np


def sparse_band(n, width, per_row, seed, weak_diag=0.2, permute=True, value_set=None):
    """Random band of the given width: per_row off-diagonal offsets per row
    from [-width/2, width/2] minus 0, plus the diagonal; lognormal values
    with random signs, a fraction of the diagonal scaled by 1e-3, then a
    symmetric random permutation.  value_set: draw values from this small
    set instead (many equal keys in the DB and RCM queues)."""
    rng = np.random.default_rng(seed)
    h = width // 2
    rows, cols = [], []
    for i in range(n):
        offs = rng.choice(np.r_[-h:0, 1:h + 1], size=min(per_row, 2 * h), replace=False)
        for o in offs:
            j = i + int(o)
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
        rows.append(i)
        cols.append(i)
    rows, cols = np.array(rows), np.array(cols)
    if value_set is None:
        vals = rng.lognormal(0.0, 1.0, rows.size) * rng.choice([-1.0, 1.0], rows.size)
    else:
        vals = rng.choice(np.asarray(value_set, dtype=float), rows.size)
    diag = rows == cols
    weak = diag & (rng.random(rows.size) < weak_diag)
    vals[weak] *= 1e-3
    if permute:
        p = rng.permutation(n)
        rows, cols = p[rows], p[cols]
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy()
