"""Learned option selector: a small network, trained on the GPU by the
library's own kernels (``Mlp``, sap_mlp_* of the C ABI), that picks SaP
options (partition count, SaP-C LU_UL / SaP-C LU_only / SaP-D, DB on/off,
drop-off fraction) for a matrix from 24 features of it, trained on solve
timings the solver itself measures.  The default options are the fallback:
when the chosen options do not reach the tolerance (the true residual
decides), the solve is repeated with the defaults.

Features (``features``): size and density, the raw bandwidth and the
distribution of |i - j|, row-length spread, the diagonal-dominance
statistics of the testDB mode (d_i = |a_ii| / sum_{j != i} |a_ij|: minimum,
1st and 10th percentiles, median, fractions above 1 and below 0.1;
sap/graph.h:468-490, :537-561), weak and zero diagonals, value magnitude
and sign statistics.  They are computed on the host from the matrix as
given (before any reordering), so choosing costs no setup.

Targets: log10 of each option set's time-to-solution (setup + solve,
relative residual 1e-10) over the default options' time on the same
matrix; option sets that do not converge get the log of 10x the matrix's
slowest converged time.  The network regresses all of them at once (masked
mean squared error), and the choice is the option set with the smallest
prediction.
"""
import math
import time

import numpy as np

N_FEATURES = 24

# option grid: (P, kind, performDB, dropOffFraction); kind 0 = SaP-C LU_UL,
# 1 = SaP-C LU_only, 2 = SaP-D (Block).  Banded inputs ignore DB and
# drop-off (they are not reordered).
PARTS = (4, 8, 16, 32, 64)
KINDS = ("spike_lu_ul", "spike_lu_only", "block")
GRID = tuple((P, kd, db, drop) for P in PARTS for kd in range(3) for db in (1, 0) for drop in (0.0, 0.02))
DEFAULT = (16, 1, 1, 0.0)   # the library defaults (Spike, LU_only, DB on, no drop-off) at P = 16
DEFAULT_INDEX = GRID.index(DEFAULT)


# ---------------------------------------------------------------- features
def _row_stats(n, rows, cols, vals):
    """Per-row |a_ii|, sum_{j != i} |a_ij|, max |a_ij|, entry counts."""
    a = np.abs(vals)
    diag = np.zeros(n)
    isd = rows == cols
    diag[rows[isd]] = a[isd]
    off = np.bincount(rows[~isd], weights=a[~isd], minlength=n)
    mx = np.zeros(n)
    np.maximum.at(mx, rows, a)
    cnt = np.bincount(rows, minlength=n).astype(np.float64)
    return diag, off, mx, cnt


def _features(n, rows, cols, vals, is_csr):
    diag, off, mx, cnt = _row_stats(n, rows, cols, vals)
    dist = np.abs(rows - cols)
    kraw = int(dist.max()) if dist.size else 0
    with np.errstate(divide="ignore", invalid="ignore"):
        d = np.where(off > 0, diag / np.where(off > 0, off, 1.0), np.inf)
    ds = np.sort(d)
    fin = np.where(np.isfinite(ds), ds, 1e8)
    nz = vals[vals != 0]
    la = np.log10(np.abs(nz)) if nz.size else np.zeros(1)
    lg = lambda x: math.log10(1e-8 + float(min(x, 1e8)))   # noqa: E731
    offv = np.abs(vals[(rows != cols) & (vals != 0)])
    dv = diag[diag > 0]
    f = [
        float(is_csr),
        math.log10(n),
        math.log10(max(vals.size, 1) / n),
        math.log10(1 + kraw),
        kraw / n,
        math.log10(1 + dist.mean()) if dist.size else 0.0,
        float(cnt.std() / max(cnt.mean(), 1e-12)),
        math.log10(max(cnt.max(), 1)),
        lg(fin[0]),
        lg(fin[n // 100]),
        lg(fin[n // 10]),
        lg(fin[n // 2]),
        float(np.mean(d >= 1.0)),
        float(np.mean(d < 0.1)),
        float(np.mean(diag < 1e-2 * mx)),
        float(la.mean()),
        float(la.std()),
        float(np.mean(vals < 0)),
        math.log10(n / (2 * kraw + 1)),
        float(np.mean(dist <= kraw / 4)) if dist.size else 1.0,
        float(np.mean(dist <= kraw / 2)) if dist.size else 1.0,
        math.log10((np.median(dv) if dv.size else 1e-8) / (np.median(offv) if offv.size else 1.0) + 1e-12),
        float(np.mean(diag == 0)),
        float(dist.mean() / max(kraw, 1)) if dist.size else 0.0,
    ]
    assert len(f) == N_FEATURES
    return np.array(f, dtype=np.float64)


def features_band(band):
    """Features of a row-major band (host numpy array n x (2k+1))."""
    n, w = band.shape
    k = (w - 1) // 2
    i = np.repeat(np.arange(n), w)
    j = i + np.tile(np.arange(-k, k + 1), n)
    v = band.reshape(-1)
    keep = (j >= 0) & (j < n) & (v != 0)
    return _features(n, i[keep], j[keep], v[keep], False)


def features_csr(rp, ci, val):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    return _features(n, rows, np.asarray(ci, dtype=np.int64), np.asarray(val, dtype=np.float64), True)


# ----------------------------------------------------------------- options
def options_for(sap, cfg, is_csr, **extra):
    P, kd, db, drop = cfg
    o = sap.Options(relTol=1e-10, refine=6, maxNumIterations=30, **extra)
    o.precondType = sap.Block if kd == 2 else sap.Spike
    o.factMethod = sap.LU_UL if kd == 0 else sap.LU_only
    if is_csr:
        o.performDB = db
        o.dropOffFraction = drop
        # LU_UL needs the constant-bandwidth path
        o.variableBandwidth = 0 if kd == 0 else 1
    return o


def unique_configs(is_csr):
    """Grid indices that give distinct solves (banded inputs: DB and
    drop-off do not apply) and, per grid index, the one it equals."""
    if is_csr:
        return list(range(len(GRID))), list(range(len(GRID)))
    rep, same = {}, []
    for g, (P, kd, db, drop) in enumerate(GRID):
        key = (P, kd)
        rep.setdefault(key, g)
        same.append(rep[key])
    return sorted(set(same)), same


def time_config(sap, torch, A, b, cfg, is_csr, reps=1):
    """Best-of-reps time (ms) of setup + solve with cfg, and whether the
    true relative residual reached 1e-10."""
    best, ok_all = None, True
    for _ in range(reps):
        s = sap.Solver(cfg[0], options_for(sap, cfg, is_csr))
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        try:
            s.setup(A)
            s.solve(None, b, x)
            torch.cuda.synchronize()
            t = (time.perf_counter() - t0) * 1e3
            r = torch.empty_like(b)
            (sap.csr_spmv if is_csr else sap.band_spmv)(A, x, r)
            ok = bool(torch.isfinite(x).all()) and float(torch.linalg.norm(b - r) / torch.linalg.norm(b)) <= 1e-10
        except sap.SapError:
            torch.cuda.synchronize()
            t, ok = (time.perf_counter() - t0) * 1e3, False
        best = t if best is None else min(best, t)
        ok_all &= ok
    return best, ok_all


# ------------------------------------------------------------------ corpus
def corpus_matrix(sap, torch, i, rng):
    """Matrix i of the corpus: banded (as the banded workloads, dominance in
    [0.5, 2]) or general sparse (random band of width 40..300 under a
    symmetric permutation, weak diagonals), with its right-hand side."""
    from . import workloads as W
    if rng.random() < 0.6:
        n = int(rng.integers(20_000, 100_001))
        k = int(rng.integers(10, 151))
        dom = float(rng.uniform(0.5, 2.0))
        A = sap.BandedMatrix.generate(n, k, dom, 7000 + i)
        f = features_band(A.band.cpu().numpy())
        xt = sap.gen_vector(n, 8000 + i)
        b = torch.empty_like(xt)
        sap.band_spmv(A, xt, b)
        desc = {"kind": "banded", "n": n, "k": k, "dominance": dom}
        return A, b, f, False, desc
    n = int(rng.integers(10_000, 30_001))
    width = int(rng.integers(40, 301))
    per_row = int(rng.integers(3, 12))
    weak = float(rng.uniform(0.0, 0.3))
    dom = float(rng.uniform(0.5, 2.0))
    rp, ci, val = W.general_sparse(n, width, per_row, seed=9000 + i, weak_seed=9500 + i, weak_diag=0.0)
    # diagonal = dominance x the row's off-diagonal abs-sum (random sign),
    # then a fraction of the diagonals shrunk by 1e-3 (rows DB must repair)
    rows = np.repeat(np.arange(n), np.diff(rp))
    isd = rows == ci
    offsum = np.bincount(rows[~isd], weights=np.abs(val[~isd]), minlength=n)
    sign = np.where(np.random.default_rng(9700 + i).random(n) < 0.5, -1.0, 1.0)
    shrink = np.where(np.random.default_rng(9800 + i).random(n) < weak, 1e-3, 1.0)
    val = val.copy()
    # (a row whose off-diagonal entries all fell outside the matrix keeps a
    # diagonal of magnitude dom)
    val[isd] = (dom * np.where(offsum > 0, offsum, 1.0) * sign * shrink)[rows[isd]]
    A = sap.CsrMatrix.from_host(rp, ci, val)
    f = features_csr(rp, ci, val)
    xt = torch.from_numpy(W.uniform_vector(n, 9900 + i)).cuda()
    b = torch.empty_like(xt)
    sap.csr_spmv(A, xt, b)
    desc = {"kind": "csr", "n": n, "width": width, "per_row": per_row, "weak": weak, "dominance": dom}
    return A, b, f, True, desc


def generate_corpus(sap, torch, count, start=0, seed=4, log=print, on_row=None):
    """Features (count x 24), times (count x |GRID|, ms) and convergence
    flags of corpus matrices start .. start + count - 1 (matrix i drawn from
    its own generator, seeded by (seed, i)), each option set timed by the
    solver."""
    F = np.zeros((count, N_FEATURES))
    T = np.zeros((count, len(GRID)))
    OK = np.zeros((count, len(GRID)), dtype=bool)
    kinds = []
    for q in range(count):
        i = start + q
        rng = np.random.default_rng((seed, i))
        A, b, f, is_csr, desc = corpus_matrix(sap, torch, i, rng)
        uniq, same = unique_configs(is_csr)
        t_u, ok_u = {}, {}
        time_config(sap, torch, A, b, GRID[uniq[0]], is_csr)   # warm-up of this size
        for g in uniq:
            t_u[g], ok_u[g] = time_config(sap, torch, A, b, GRID[g], is_csr, reps=1 if is_csr else 2)
        for g in range(len(GRID)):
            T[q, g], OK[q, g] = t_u[same[g]], ok_u[same[g]]
        F[q] = f
        kinds.append(1 if is_csr else 0)
        log(f"[corpus] {i} ({q + 1}/{count}) {desc} default {T[q, DEFAULT_INDEX]:.1f} ms "
            f"ok={OK[q, DEFAULT_INDEX]} best {T[q][OK[q]].min() if OK[q].any() else float('nan'):.1f} ms")
        if on_row is not None:
            on_row(q + 1, F[:q + 1], T[:q + 1], OK[:q + 1], np.array(kinds))
    return F, T, OK, np.array(kinds)


# ---------------------------------------------------------------- training
def targets(T, OK):
    """log10(time / default time) per option set; non-converged option sets
    get log10 of 10x the matrix's slowest converged time (relative to its
    default).  Matrices where nothing converged are masked out."""
    Y = np.zeros_like(T)
    M = np.ones_like(T)
    for i in range(T.shape[0]):
        if not OK[i].any():
            M[i] = 0.0
            continue
        ref = T[i, DEFAULT_INDEX] if OK[i, DEFAULT_INDEX] else T[i][OK[i]].max() * 10
        worst = T[i][OK[i]].max() * 10
        Y[i] = np.log10(np.where(OK[i], T[i], worst) / ref)
    return Y, M


class Selector:
    """Feature standardisation + an ensemble of networks (differently
    initialised, the predictions averaged: five-fold cross-validation on the
    committed corpus gave 97.9% of matrices no slower than the defaults with
    five networks against 95.0% with one); ``choose`` returns a grid index."""

    def __init__(self, sap, torch, hidden=64, seed=0, models=5):
        self.sap, self.torch = sap, torch
        self.nets = [sap.Mlp(N_FEATURES, len(GRID), hidden, seed=seed + 101 * j) for j in range(models)]
        self.mu = np.zeros(N_FEATURES)
        self.sd = np.ones(N_FEATURES)

    def _x(self, F):
        X = (np.atleast_2d(F) - self.mu) / self.sd
        return self.torch.from_numpy(X.astype(np.float32)).cuda()

    def fit(self, F, T, OK, epochs=3000, lr=2e-3):
        """Trains every network on the GPU; returns the first one's losses."""
        self.mu = F.mean(axis=0)
        self.sd = F.std(axis=0) + 1e-6
        Y, M = targets(T, OK)
        f32 = lambda a: self.torch.from_numpy(a.astype(np.float32)).cuda()   # noqa: E731
        X, Yd, Md = self._x(F), f32(Y), f32(M)
        losses = [net.train(X, Yd, Md, epochs, lr=lr) for net in self.nets]
        return losses[0]

    def predict(self, F):
        X = self._x(F)
        return np.mean([net.forward(X).cpu().numpy() for net in self.nets], axis=0)

    def choose(self, F):
        return np.argmin(self.predict(F), axis=1)


def evaluate(choice, T, OK):
    """Per matrix: time of the chosen option set, with the default's time
    added when the choice did not converge (the fallback re-solves with the
    defaults), over the default's time."""
    out = []
    for i, g in enumerate(choice):
        t = T[i, g] if OK[i, g] else T[i, g] + T[i, DEFAULT_INDEX]
        out.append(t / T[i, DEFAULT_INDEX])
    return np.array(out)
