"""Dense numpy restatements used to pin the oracle (CPU, small sizes only).

The truncated-SPIKE preconditioner is built here from its definition
(spikes V_i = A_i^{-1}[0; B_i], W_{i+1} = A_{i+1}^{-1}[C_{i+1}; 0], their
K x K tips, the reduced blocks R_i = [[I, V_i], [W_{i+1}, I]], purification
and a second block-diagonal solve), independently of the factor-and-sweep
formulation that the oracle and the HIP path use.
"""
import numpy as np


def dense_from_band(band):
    n, w = band.shape
    k = (w - 1) // 2
    A = np.zeros((n, n))
    for i in range(n):
        for d in range(w):
            j = i + d - k
            if 0 <= j < n:
                A[i, j] = band[i, d]
    return A


def band_from_dense(A, k):
    n = A.shape[0]
    band = np.zeros((n, 2 * k + 1))
    for i in range(n):
        for d in range(2 * k + 1):
            j = i + d - k
            if 0 <= j < n:
                band[i, d] = A[i, j]
    return band


def clamp_partitions(n, k, P):
    return min(P, max(1, n // max(k + 1, 2 * k)))


def partitions(n, P):
    q, r = divmod(n, P)
    out, s = [], 0
    for p in range(P):
        ln = q + 1 if p < r else q
        out.append((s, ln))
        s += ln
    return out


def block_solve(A, parts, v):
    out = np.empty_like(v)
    for s, ln in parts:
        out[s:s + ln] = np.linalg.solve(A[s:s + ln, s:s + ln], v[s:s + ln])
    return out


def spike_tips(A, k, parts):
    tips = []
    for i in range(len(parts) - 1):
        s0, l0 = parts[i]
        s1, l1 = parts[i + 1]
        e0 = s0 + l0
        Bi = A[e0 - k:e0, s1:s1 + k]
        Ci = A[s1:s1 + k, e0 - k:e0]
        rhs_v = np.zeros((l0, k))
        rhs_v[-k:] = Bi
        rhs_w = np.zeros((l1, k))
        rhs_w[:k] = Ci
        V = np.linalg.solve(A[s0:e0, s0:e0], rhs_v)[-k:]
        W = np.linalg.solve(A[s1:s1 + l1, s1:s1 + l1], rhs_w)[:k]
        tips.append((V, W, Bi, Ci))
    return tips


def spike_apply(A, k, P, f):
    """M^{-1} f for the truncated SPIKE (SaP-C) preconditioner."""
    n = A.shape[0]
    parts = partitions(n, P)
    g = block_solve(A, parts, f)
    if P == 1:
        return g
    tips = spike_tips(A, k, parts)
    xi = g.copy()
    for i, (V, W, _, _) in enumerate(tips):
        s1 = parts[i + 1][0]
        R = np.block([[np.eye(k), V], [W, np.eye(k)]])
        z = np.linalg.solve(R, np.concatenate([g[s1 - k:s1], g[s1:s1 + k]]))
        xi[s1 - k:s1] = z[:k]
        xi[s1:s1 + k] = z[k:]
    sol = f.copy()
    for i, (_, _, Bi, Ci) in enumerate(tips):
        s1 = parts[i + 1][0]
        sol[s1 - k:s1] -= Bi @ xi[s1:s1 + k]
        sol[s1:s1 + k] -= Ci @ xi[s1 - k:s1]
    return block_solve(A, parts, sol)


def block_apply(A, k, P, f):
    """M^{-1} f for the block-diagonal (SaP-D) preconditioner."""
    return block_solve(A, partitions(A.shape[0], P), f)


def unpack_colband(fac, k):
    """Column-major factor band (n, 2k+1) -> dense matrix with the stored entries."""
    n = fac.shape[0]
    M = np.zeros((n, n))
    for c in range(n):
        for d in range(2 * k + 1):
            r = c + d - k
            if 0 <= r < n:
                M[r, c] = fac[c, d]
    return M
