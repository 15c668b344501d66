"""Host reordering for CSR input (DB, RCM, drop-off): the oracle pinned by
closed forms and library routines, then the shipped host code
(sap_reorder_csr in libsap_amd.so, no GPU call) bit-exact against the oracle.

What pins the oracle:
  * its priority queue against libstdc++'s std::priority_queue, compiled here
    with g++, on push/pop sequences full of equal keys (the tie-breaking the
    reference inherits from libstdc++);
  * DB: the matching maximises prod |a_ij| (scipy's linear_sum_assignment
    on -log|a| gives the optimum), and the scaled matrix has |a| <= 1 with
    ones on the matched entries (the dual feasibility MC64 scaling promises);
  * RCM: a permutation, the bandwidth it reports is that of the permuted
    matrix, a randomly permuted path graph comes back with bandwidth 1;
  * drop-off: the kept entries are exactly those within the reported
    half-bandwidth, the dropped weight does not exceed the requested
    fraction, and dropping the next diagonal would.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import graph_oracle as G
from tests.csr_cases import sparse_band

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def sap():
    import saplibrary_amd
    if not os.path.exists(saplibrary_amd.LIB_PATH):
        pytest.skip("libsap_amd.so not built")
    saplibrary_amd.lib()
    return saplibrary_amd


# ------------------------------------------------------------ priority queue
PQ_SRC = r"""
#include <cstdio>
#include <queue>
#include <vector>
struct NK { int node; long key; };
struct Cmp { bool operator()(const NK& a, const NK& b) const { return a.key > b.key; } };
int main() {
    std::priority_queue<NK, std::vector<NK>, Cmp> q;
    char op; int node; long key;
    while (std::scanf(" %c", &op) == 1) {
        if (op == 'u') { std::scanf("%d %ld", &node, &key); q.push({node, key}); }
        else { std::printf("%d %ld\n", q.top().node, q.top().key); q.pop(); }
    }
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_priority_queue_matches_libstdcxx(tmp_path):
    src = tmp_path / "pq.cpp"
    src.write_text(PQ_SRC)
    exe = tmp_path / "pq"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    rng = np.random.default_rng(5)
    for trial in range(20):
        ops, q, expect = [], G.LibstdcxxPQ(), []
        size = 0
        for step in range(400):
            if size and rng.random() < 0.45:
                ops.append("o")
                expect.append(q.top())
                q.pop()
                size -= 1
            else:
                node, key = int(rng.integers(0, 1000)), int(rng.integers(0, 4 + trial))
                ops.append(f"u {node} {key}")
                q.push(node, key)
                size += 1
        while size:
            ops.append("o")
            expect.append(q.top())
            q.pop()
            size -= 1
        out = subprocess.run([str(exe)], input="\n".join(ops), capture_output=True, text=True, check=True).stdout
        got = [tuple(int(x) for x in line.split()) for line in out.strip().splitlines()]
        assert got == [(a, int(b)) for a, b in expect], f"trial {trial}"


def test_priority_queue_pops_in_key_order():
    q = G.LibstdcxxPQ()
    rng = np.random.default_rng(1)
    keys = rng.random(300)
    for i, k in enumerate(keys):
        q.push(i, float(k))
    out = []
    while not q.empty():
        out.append(q.top()[1])
        q.pop()
    assert out == sorted(keys.tolist())


# ------------------------------------------------------------------------ DB
def _dense(rp, ci, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            A[i, ci[e]] = val[e]
    return A


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_db_matching_is_maximum_product(seed):
    from scipy.optimize import linear_sum_assignment
    n = 120
    rp, ci, val = sparse_band(n, 30, 5, seed)
    try:
        r = G.reorder(rp, ci, val, do_db=True, scale=True, max_bw=64, prec="double")
    except ValueError as e:
        pytest.skip(f"DB raised {e} on this input")
    A = np.abs(_dense(rp, ci, val, n))
    # the row of original row i is db_perm[i]; its matched column is db_perm[i]
    cols = r["db_perm"]
    assert sorted(cols.tolist()) == list(range(n))
    assert np.all(A[np.arange(n), cols] > 0)
    got = np.sum(np.log(A[np.arange(n), cols]))
    C = np.where(A > 0, -np.log(np.where(A > 0, A, 1.0)), 1e6)
    ri, cj = linear_sum_assignment(C)
    best = np.sum(np.log(A[ri, cj]))
    assert got == pytest.approx(best, abs=1e-9 * n)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_db_scaling_is_dual_feasible(seed):
    n = 150
    rp, ci, val = sparse_band(n, 24, 4, seed + 10)
    try:
        r = G.reorder(rp, ci, val, do_db=True, scale=True, prec="double")
    except ValueError as e:
        pytest.skip(f"DB raised {e} on this input")
    # the reordered matrix: rows by row_perm (DB then RCM), columns by RCM
    S = _dense(r["rp"], r["ci"], r["val"], n)
    assert np.max(np.abs(S)) <= 1 + 1e-12
    assert np.allclose(np.abs(np.diag(S)), 1.0, rtol=0, atol=1e-12)
    # the scalings reproduce it: S[row_perm[i], colperm[j]] = r_i a_ij c_j
    A = _dense(rp, ci, val, n)
    D = r["row_scale"][:, None] * A * r["col_scale"][None, :]
    colperm = np.empty(n, dtype=np.int64)
    colperm[r["col_reord"]] = np.arange(n)
    X = np.zeros((n, n))
    X[np.ix_(r["row_perm"], colperm)] = D
    assert np.allclose(X, S, rtol=1e-14, atol=0)


def test_db_singular_matrix_raises():
    # column 1 is empty: no perfect matching
    rp = np.array([0, 2, 3, 4], dtype=np.int32)
    ci = np.array([0, 2, 0, 2], dtype=np.int32)
    val = np.array([1.0, 2.0, 3.0, 4.0])
    with pytest.raises(ValueError, match="Matrix_singular"):
        G.reorder(rp, ci, val, prec="double")


# ----------------------------------------------------------------------- RCM
def test_rcm_recovers_a_permuted_path():
    n = 300
    rng = np.random.default_rng(3)
    p = rng.permutation(n)
    rows = np.r_[np.arange(n), np.arange(n - 1), np.arange(1, n)]
    cols = np.r_[np.arange(n), np.arange(1, n), np.arange(n - 1)]
    import scipy.sparse as sp
    A = sp.csr_matrix((np.r_[np.full(n, 4.0), -np.ones(2 * n - 2)], (p[rows], p[cols])), shape=(n, n))
    A.sort_indices()
    r = G.reorder(A.indptr, A.indices, A.data, do_db=False, prec="double")
    assert r["k_db"] > 100 and r["k_reorder"] == 1 and r["k"] == 1


@pytest.mark.parametrize("seed", [4, 5])
def test_rcm_reports_the_bandwidth_of_its_permutation(seed):
    n = 400
    rp, ci, val = sparse_band(n, 40, 4, seed)
    r = G.reorder(rp, ci, val, do_db=False, prec="double")
    perm = np.empty(n, dtype=np.int64)
    perm[r["col_reord"]] = np.arange(n)
    assert sorted(r["col_reord"].tolist()) == list(range(n))
    assert np.array_equal(r["row_perm"], perm)          # without DB the rows follow RCM alone
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert r["k_reorder"] == np.max(np.abs(perm[rows] - perm[ci]))
    assert r["k_reorder"] <= r["k_db"]
    assert r["k_reorder"] <= 60


# ------------------------------------------------------------------ drop-off
@pytest.mark.parametrize("frac", [0.001, 0.01, 0.05])
def test_drop_off_keeps_the_band_and_respects_the_budget(frac):
    n = 300
    rp, ci, val = sparse_band(n, 60, 6, 7)
    full = G.reorder(rp, ci, val, do_db=False, prec="double")
    r = G.reorder(rp, ci, val, do_db=False, drop_off=frac, prec="double")
    k = r["k"]
    assert k <= full["k_reorder"]
    rows = np.repeat(np.arange(n), np.diff(full["rp"]))
    d = np.abs(rows - full["ci"])
    w = full["val"] ** 2
    rows2 = np.repeat(np.arange(n), np.diff(r["rp"]))
    kept = np.max(np.abs(rows2 - r["ci"]))
    # whole diagonals go: the kept entries are all those within the widest
    # kept diagonal.  The reported half-bandwidth is that diagonal, or the
    # last diagonal removed when the loop removed any (sap/graph.h:700-707)
    assert r["ci"].size == np.count_nonzero(d <= kept)
    assert k >= kept and (k == kept or np.count_nonzero(d == k) > 0)
    assert np.all(d[(d > kept) & (d <= k)] == k) or k == kept
    dropped = w[d > kept].sum() / w.sum()
    assert dropped <= frac * (1 + 1e-12)
    assert r["dropoff"] == pytest.approx(dropped, rel=1e-9, abs=1e-15)
    if kept > 0:
        assert w[d >= kept].sum() / w.sum() > frac        # one more diagonal is too much


def test_max_bandwidth_caps_the_band():
    n = 200
    rp, ci, val = sparse_band(n, 50, 5, 9)
    r = G.reorder(rp, ci, val, do_db=False, max_bw=8, prec="double")
    rows = np.repeat(np.arange(n), np.diff(r["rp"]))
    assert r["k"] <= 8 and np.max(np.abs(rows - r["ci"])) <= r["k"]
    # maxBandwidth <= 64 switches RCM off (sap/precond.h:1737-1738)
    assert np.array_equal(r["col_reord"], np.arange(n))


# ------------------------------------------------- shipped code vs the oracle
CASES = [
    # n, width, per_row, seed, options
    (200, 30, 5, 11, dict()),
    (300, 50, 6, 12, dict(drop_off=0.01)),
    (250, 40, 4, 13, dict(scale=False)),
    (180, 20, 3, 14, dict(do_db=False)),
    (220, 36, 5, 15, dict(first_stage_only=True)),
    (260, 44, 5, 16, dict(max_bw=12)),
    (240, 30, 4, 17, dict(value_set=[1.0, -1.0, 2.0, 0.5])),
    (150, 400, 3, 18, dict(drop_off=0.02)),
    (600, 80, 8, 19, dict(drop_off=0.005)),
]


def _options(sap, o, prec):
    opts = sap.Options()
    opts.performDB = int(o.get("do_db", True))
    opts.applyScaling = int(o.get("scale", True))
    opts.dbFirstStageOnly = int(o.get("first_stage_only", False))
    opts.maxBandwidth = int(o.get("max_bw", 2 ** 31 - 1))
    opts.dropOffFraction = float(o.get("drop_off", 0.0))
    opts.precValueType = sap.PREC_FLOAT if prec == "float" else sap.PREC_DOUBLE
    return opts


@pytest.mark.parametrize("prec", ["float", "double"])
@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}w{c[1]}-{c[3]}" for c in CASES])
def test_product_reordering_is_bit_exact(sap, case, prec):
    n, width, per_row, seed, o = case
    o = dict(o)
    vs = o.pop("value_set", None)
    rp, ci, val = sparse_band(n, width, per_row, seed, value_set=vs)
    try:
        ref = G.reorder(rp, ci, val, prec=prec, **o)
    except ValueError as e:
        code = {"Matrix_singular": -5, "Negative_DB_weight": -2}[str(e)]
        with pytest.raises(sap.SapError) as ei:
            sap.reorder_csr(rp, ci, val, _options(sap, o, prec))
        assert ei.value.reason == code
        return
    got = sap.reorder_csr(rp, ci, val, _options(sap, o, prec))
    for f in ("k", "k_db", "k_reorder"):
        assert got[f] == ref[f], f
    assert got["dropoff"] == ref["dropoff"]
    for f in ("db_perm", "row_perm", "col_reord", "rp", "ci"):
        assert np.array_equal(got[f], ref[f]), f
    for f in ("row_scale", "col_scale", "val"):
        assert np.array_equal(got[f], ref[f]), f


def test_product_reordering_rejects_singular(sap):
    rp = np.array([0, 2, 3, 4], dtype=np.int32)
    ci = np.array([0, 2, 0, 2], dtype=np.int32)
    val = np.array([1.0, 2.0, 3.0, 4.0])
    with pytest.raises(sap.SapError) as ei:
        sap.reorder_csr(rp, ci, val)
    assert ei.value.reason == -5
    # no entries at all
    with pytest.raises(sap.SapError) as ei:
        sap.reorder_csr(np.zeros(4, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0))
    assert ei.value.reason == -5


# ------------------------------------------------ variable-bandwidth assembly
VAR_CASES = [(2000, 60, 6, 51, 8), (1500, 40, 5, 52, 5), (1200, 90, 7, 53, 3), (900, 30, 4, 54, 16)]


def _var_ref(n, width, per_row, seed, P, prec="float"):
    rp, ci, val = sparse_band(n, width, per_row, seed)
    ref = G.reorder(rp, ci, val, prec=prec)
    Pc = min(P, max(n // (ref["k"] + 1), 1))
    return rp, ci, val, ref, Pc, G.assemble_var(n, ref["k"], Pc, ref["rp"], ref["ci"], ref["val"], prec=prec)


@pytest.mark.parametrize("case", VAR_CASES, ids=[f"n{c[0]}P{c[4]}" for c in VAR_CASES])
def test_var_assembly_invariants(case):
    n, width, per_row, seed, P = case
    rp, ci, val, ref, P, v = _var_ref(n, width, per_row, seed, P, prec="double")
    k = ref["k"]
    part, rem = n // P, n % P
    sizes = np.array([part + (1 if i < rem else 0) for i in range(P)])
    starts = np.r_[0, np.cumsum(sizes)]
    # band offsets: prefix sums of n_p (2 k_p + 1)
    assert np.array_equal(v["boffsets"], np.r_[0, np.cumsum(sizes * (2 * v["ks"].astype(np.int64) + 1))])
    # the second-level permutation maps every partition onto itself
    sp, sr = v["second_perm"], v["second_reorder"]
    assert np.array_equal(sp[sr], np.arange(n))
    pof = np.searchsorted(starts, np.arange(n), side="right") - 1
    assert np.array_equal(pof[sp], pof)
    # every entry of the reordered matrix is either in a diagonal block (at
    # its permuted place) or in a coupling block (as it is), nowhere else
    rows = np.repeat(np.arange(n), np.diff(ref["rp"]))
    cols = ref["ci"]
    same = pof[rows] == pof[cols]
    D = {}
    drows = np.repeat(np.arange(n), np.diff(v["drp"]))
    for r, c, x in zip(drows, v["dci"], v["dval"]):
        D[(int(r), int(c))] = x
    assert len(D) == int(same.sum())
    for r, c, x in zip(rows[same], cols[same], ref["val"][same]):
        assert D[(int(sp[r]), int(sp[c]))] == x
    kk = k * k
    nrb = 0
    for r, c, x in zip(rows[~same], cols[~same], ref["val"][~same]):
        pr, pc = pof[r], pof[c]
        if pc > pr:   # B_i: bottom rows of partition pr, first columns of pc
            idx = pr * 2 * kk + (r + k - starts[pr + 1]) * k + (c - starts[pc])
        else:         # C_{i+1}: top rows of partition pr, last columns of pc
            idx = (pc * 2 + 1) * kk + (r - starts[pr]) * k + (c - starts[pc + 1] + k)
        assert v["offdiags"][idx] == x
        nrb += 1
    assert np.count_nonzero(v["offdiags"]) == nrb
    # WV holds the same blocks with their used columns packed: right blocks
    # from the left, left blocks from the right, perms giving the column
    for i in range(P - 1):
        for side, widths, perms in ((0, v["widths_right"], v["perms_right"]), (1, v["widths_left"], v["perms_left"])):
            base = (2 * i + side) * kk
            Off = v["offdiags"][base:base + kk].reshape(k, k)           # [row][col]
            W = v["wv"][base:base + kk].reshape(k, k)                    # [slot][row]
            used = np.nonzero(np.any(Off != 0, axis=0))[0]
            assert widths[i] == used.size
            slots = range(widths[i]) if side == 0 else range(k - widths[i], k)
            pc = perms[i * k:(i + 1) * k]
            assert sorted(int(pc[s]) for s in slots) == sorted(used.tolist())
            for s in slots:
                assert np.array_equal(W[s], Off[:, pc[s]])
    # half-bandwidth of each diagonal block, and the per-partition RCM never
    # widens it
    for p in range(P):
        m = (pof[drows] == p)
        assert v["ks"][p] == (np.max(np.abs(drows[m] - v["dci"][m])) if m.any() else 0)
        m0 = same & (pof[rows] == p)
        assert v["ks"][p] <= (np.max(np.abs(rows[m0] - cols[m0])) if m0.any() else 0)
    # first_rows: no later than the partition end, and at the smallest
    # permuted row among the last k rows with a coupling entry to the right
    exists = np.zeros(n, dtype=bool)
    exists[rows[~same & (pof[cols] > pof[rows])]] = True
    for i in range(P - 1):
        lr = starts[i + 1]
        cand = [int(sp[j]) for j in range(lr - k, lr) if exists[j]]
        assert v["first_rows"][i] == min([lr] + cand)


@pytest.mark.parametrize("prec", ["float", "double"])
@pytest.mark.parametrize("case", VAR_CASES, ids=[f"n{c[0]}P{c[4]}" for c in VAR_CASES])
def test_product_var_assembly_is_bit_exact(sap, case, prec):
    n, width, per_row, seed, P = case
    rp, ci, val, ref, Pc, v = _var_ref(n, width, per_row, seed, P, prec=prec)
    got = sap.reorder_csr(rp, ci, val, _options(sap, {}, prec), numPartitions=P)
    assert got["P"] == Pc
    for f, x in v.items():
        assert np.array_equal(got["var"][f], x), f


# ------------------------------------------------------------------- testDB
def _dominance(n, rp, ci, val):
    d = np.empty(n)
    for i in range(n):
        cols, v = ci[rp[i]:rp[i + 1]], np.abs(val[rp[i]:rp[i + 1]].astype(np.float64))
        diag = v[cols == i].sum()
        d[i] = diag / v[cols != i].sum()
    d.sort()
    return d[0], d[n // 100]


@pytest.mark.parametrize("prec", ["float", "double"])
def test_testdb_statistics(sap, prec):
    """testDB (sap/graph.h:468-490, :537-561): the dominance statistics of
    the input and of the DB-permuted and scaled matrix, computed directly;
    the test stops after DB (no RCM: the bandwidth is the one after DB), and
    after DB scaling every matched entry is 1 in magnitude and the others at
    most 1, so the minimum ratio cannot drop below that of |a_ii| = 1 over
    the row's other entries."""
    n = 3000
    rp, ci, val = sparse_band(n, 60, 6, 81)
    T = np.float32 if prec == "float" else np.float64
    vT = val.astype(T)
    opts = sap.Options(testDB=1, precValueType=sap.PREC_FLOAT if prec == "float" else sap.PREC_DOUBLE)
    r = sap.reorder_csr(rp, ci, val, opts)
    dmin, dp1 = _dominance(n, rp, ci, vT)
    assert r["diag_dom_ori"] == dmin and r["d_p1_ori"] == dp1
    dmin2, dp12 = _dominance(n, r["rp"], r["ci"], r["val"].astype(T))
    assert r["diag_dom"] == dmin2 and r["d_p1"] == dp12
    assert r["k"] == r["k_db"] == r["k_reorder"]
    assert np.array_equal(r["col_reord"], np.arange(n))
    rows = np.repeat(np.arange(n), np.diff(r["rp"]))
    assert np.abs(r["val"][rows == r["ci"]]).min() > 0.999
    # |a_ii| = 1 and the others at most 1: d_i >= 1 / (entries of row i - 1)
    others = np.diff(r["rp"]) - 1
    assert r["diag_dom"] >= (1.0 / others.max()) * (1 - 1e-5)
    assert r["d_p1"] > r["d_p1_ori"]
