"""Parity of the CSR path (through the C ABI) with the oracle.

  * CSR SpMV: FP64, within 1e-13 of the largest |y| (scipy's product, a
    different summation order).
  * CSR setup without reordering: the band it builds is the band of the
    matrix, so factors and M^{-1} v are bit-identical to the banded setup.
  * CSR setup with DB / RCM / drop-off: the statistics (bandwidths, drop-off
    fraction, partitions) equal the oracle's, and M^{-1} v equals the
    oracle's chain (its reordering, its banded preconditioner on the
    reordered band, the permutations and scalings in the preconditioner
    type) within 1e-10 of the largest |z| in FP64; in FP32 within 3e-5 of
    it or 4x the oracle's own FP32 error, whichever is larger (these
    matrices are far from diagonally dominant, so FP32 rounding is amplified
    by the conditioning of the diagonal blocks).
  * End to end: FP64 residual of the solution on a matrix of the sparse
    workload's kind (random band of width 300, ~12 entries per row, weak
    diagonals, symmetric random permutation).
"""
import numpy as np
import pytest

from oracle import graph_oracle as G
from tests.csr_cases import sparse_band

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sap():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import saplibrary_amd
    saplibrary_amd.lib()
    return saplibrary_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def csr_of_band(band):
    n, w = band.shape
    k = (w - 1) // 2
    rp, ci, val = [0], [], []
    for i in range(n):
        for j in range(max(0, i - k), min(n, i + k + 1)):
            ci.append(j)
            val.append(band[i, k + j - i])
        rp.append(len(ci))
    return np.array(rp, dtype=np.int32), np.array(ci, dtype=np.int32), np.array(val)


@pytest.mark.parametrize("n,per_row", [(5000, 3), (4000, 12), (3000, 40), (7, 2)])
def test_csr_spmv(sap, n, per_row):
    import scipy.sparse as sp
    rp, ci, val = sparse_band(n, max(4, 3 * per_row), per_row, 1)
    A = sap.CsrMatrix.from_host(rp, ci, val)
    x = np.random.default_rng(2).uniform(-1, 1, n)
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    sap.csr_spmv(A, dev(x), y)
    ref = sp.csr_matrix((val, ci, rp), shape=(n, n)) @ x
    np.testing.assert_allclose(host(y), ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("n,k,P,fact,prec", [(3000, 20, 4, "LU_UL", 0), (6000, 80, 4, "LU_UL", 0),
                                              (4000, 50, 5, "LU_only", 0), (500, 3, 1, "LU_UL", 0),
                                              (5000, 60, 4, "LU_UL", 1), (3000, 25, 3, "LU_only", 1)])
def test_csr_without_reordering_is_the_banded_setup(sap, oracle, n, k, P, fact, prec):
    """The band converted from the row-major input (column tiles through
    LDS) and the band scattered from the CSR entries are the same band: the
    factors and M^{-1} v agree bit for bit."""
    band = oracle.gen_banded(n, k, 0.9, 41)
    rp, ci, val = csr_of_band(band)
    f = sap.LU_UL if fact == "LU_UL" else sap.LU_only
    opts = sap.Options(factMethod=f, performReorder=0, performDB=0, applyScaling=0, variableBandwidth=0,
                       precValueType=prec)
    g_csr = sap.Solver(P, opts)
    g_csr.setup(sap.CsrMatrix.from_host(rp, ci, val))
    g_band = sap.Solver(P, sap.Options(factMethod=f, precValueType=prec))
    g_band.setup(sap.BandedMatrix(dev(band)))
    st = g_csr.getStats()
    assert st.bandwidth == k and st.bandwidthDB == 0 and st.bandwidthReorder == 0
    assert g_csr.numPartitions == g_band.numPartitions
    np.testing.assert_array_equal(g_csr.factors(), g_band.factors())
    v = dev(oracle.gen_vector(n, 42))
    z1 = torch.empty(n, dtype=torch.float64, device="cuda")
    z2 = torch.empty_like(z1)
    g_csr.precond_apply(v, z1)
    g_band.precond_apply(v, z2)
    np.testing.assert_array_equal(host(z1), host(z2))


CASES = [
    # n, width, per_row, seed, P, fact, reorder options
    (1500, 60, 6, 51, 4, "LU_UL", dict()),
    (1200, 40, 5, 52, 3, "LU_only", dict(drop_off=0.01)),
    (1000, 50, 4, 53, 1, "LU_UL", dict(scale=False)),
    (1400, 30, 5, 54, 5, "LU_UL", dict(do_db=False)),
    (1300, 80, 6, 55, 2, "LU_UL", dict(max_bw=20)),
]


def _opts(sap, o, fact, prec):
    return sap.Options(
        factMethod=sap.LU_UL if fact == "LU_UL" else sap.LU_only, variableBandwidth=0,
        performDB=int(o.get("do_db", True)), applyScaling=int(o.get("scale", True)),
        maxBandwidth=int(o.get("max_bw", 2 ** 31 - 1)), dropOffFraction=float(o.get("drop_off", 0.0)),
        precValueType=sap.PREC_FLOAT if prec == "float" else sap.PREC_DOUBLE)


@pytest.mark.parametrize("prec", ["float", "double"])
@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}-{c[3]}" for c in CASES])
def test_csr_reordered_apply_matches_oracle(sap, oracle, case, prec):
    n, width, per_row, seed, P, fact, o = case
    rp, ci, val = sparse_band(n, width, per_row, seed)
    ref = G.reorder(rp, ci, val, prec=prec, **o)
    T = np.float32 if prec == "float" else np.float64
    k = ref["k"]
    band = G.csr_to_band(n, k, ref["rp"], ref["ci"], ref["val"])
    cpu = oracle.Solver(P, fact=fact, prec_double=(prec == "double"))
    assert cpu.setup(band) == 0

    g = sap.Solver(P, _opts(sap, o, fact, prec))
    g.setup(sap.CsrMatrix.from_host(rp, ci, val))
    st = g.getStats()
    assert (st.bandwidth, st.bandwidthDB, st.bandwidthReorder) == (k, ref["k_db"], ref["k_reorder"])
    assert st.actualDropOff == ref["dropoff"]
    assert g.numPartitions == cpu.num_partitions

    v = oracle.gen_vector(n, seed + 100)
    scale = o.get("scale", True)

    def chain(solver, T):
        w = np.zeros(n)
        x = v.astype(T) * (ref["row_scale"].astype(T) if scale else T(1))
        w[ref["row_perm"]] = x.astype(np.float64)
        zr = solver.apply(w)
        out = np.zeros(n)
        y = zr.astype(T) * (ref["col_scale"][ref["col_reord"]].astype(T) if scale else T(1))
        out[ref["col_reord"]] = y.astype(np.float64)
        return out

    expect = chain(cpu, T)
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    g.precond_apply(dev(v), z)
    if prec == "double":
        atol = 1e-10 * np.abs(expect).max()
    else:
        # FP32 rounding, amplified by the conditioning of the diagonal
        # blocks: the GPU may differ from the oracle by 3e-5 of max |z|, or
        # by 4x the oracle's own FP32 error (against the same preconditioner
        # applied in FP64), whichever is larger
        cpu64 = oracle.Solver(P, fact=fact, prec_double=True)
        assert cpu64.setup(band) == 0
        err32 = np.abs(expect - chain(cpu64, np.float64)).max()
        atol = max(3e-5 * np.abs(expect).max(), 4 * err32)
    np.testing.assert_allclose(host(z), expect, rtol=0, atol=atol)


def test_csr_solve_sparse_workload(sap):
    import scipy.sparse as sp
    n = 30000
    rp, ci, val = sparse_band(n, 300, 11, 61)
    A = sp.csr_matrix((val, ci, rp), shape=(n, n))
    xt = np.random.default_rng(62).uniform(-1, 1, n)
    b = A @ xt
    # (two partitions: this matrix is far from diagonally dominant and its
    # reordered band is wide, so truncated SPIKE stops converging beyond
    # P = 2 -- the oracle's SaP behaves the same on the reordered band)
    opts = sap.Options(relTol=1e-10, refine=8)      # reordering, DB, scaling, variable bandwidth
    g = sap.Solver(2, opts)
    g.setup(sap.CsrMatrix.from_host(rp, ci, val))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert g.solve(None, dev(b), x), g.getMonitorMessage()
    xs = host(x)
    assert np.linalg.norm(b - A @ xs) / np.linalg.norm(b) <= 1e-10
    st = g.getStats()
    assert st.bandwidthDB > st.bandwidthReorder >= st.bandwidth > 0
    assert st.numSpmv > 0 and g.numPartitions == 2


@pytest.mark.parametrize("prec", ["float", "double"])
def test_csr_variable_bandwidth_matches_oracle(sap, oracle, prec):
    """Default options (variableBandwidth, LU_only): the truncated-SPIKE
    operator of the reordered matrix, so the oracle's LU_only preconditioner
    on the reordered band within the tolerances above; nuKf from the
    variable-bandwidth band extents (sap/solver.h:358-394)."""
    n, P = 1500, 4
    rp, ci, val = sparse_band(n, 60, 6, 56)
    ref = G.reorder(rp, ci, val, prec=prec)
    T = np.float32 if prec == "float" else np.float64
    k = ref["k"]
    band = G.csr_to_band(n, k, ref["rp"], ref["ci"], ref["val"])
    cpu = oracle.Solver(P, fact="LU_only", prec_double=(prec == "double"))
    assert cpu.setup(band) == 0
    g = sap.Solver(P, sap.Options(precValueType=sap.PREC_FLOAT if prec == "float" else sap.PREC_DOUBLE))
    g.setup(sap.CsrMatrix.from_host(rp, ci, val))
    st = g.getStats()
    assert st.bandwidth == k and g.numPartitions == cpu.num_partitions
    Pv = min(P, max(n // (k + 1), 1))
    var = G.assemble_var(n, k, Pv, ref["rp"], ref["ci"], ref["val"], prec=prec)
    nukf = (2.0 * k * n - (var["ks_row"].sum() + var["ks_col"].sum())) / (2.0 * k * n)
    assert st.nuKf == pytest.approx(nukf, rel=1e-12)
    v = oracle.gen_vector(n, 57)
    w = np.zeros(n)
    w[ref["row_perm"]] = (v.astype(T) * ref["row_scale"].astype(T)).astype(np.float64)
    zr = cpu.apply(w)
    expect = np.zeros(n)
    expect[ref["col_reord"]] = (zr.astype(T) * ref["col_scale"][ref["col_reord"]].astype(T)).astype(np.float64)
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    g.precond_apply(dev(v), z)
    if prec == "double":
        atol = 1e-10 * np.abs(expect).max()
    else:
        cpu64 = oracle.Solver(P, fact="LU_only", prec_double=True)
        assert cpu64.setup(band) == 0
        w64 = np.zeros(n)
        w64[ref["row_perm"]] = v * ref["row_scale"]
        z64 = np.zeros(n)
        z64[ref["col_reord"]] = cpu64.apply(w64) * ref["col_scale"][ref["col_reord"]]
        atol = max(3e-5 * np.abs(expect).max(), 4 * np.abs(expect - z64).max())
    np.testing.assert_allclose(host(z), expect, rtol=0, atol=atol)


@pytest.mark.parametrize("variable", [True, False])
def test_reordering_path_partition_clamp(sap, oracle, variable):
    """One partition count on the reordering path: a request above
    n / max(k+1, 2k) is clamped to it (the reference clamps with n / (k+1)
    there, sap/precond.h:1777-1778, which allows partitions shorter than
    2k rows whose interface windows overlap), the statistics (numPartitions,
    nuKf from the variable-bandwidth extents) describe that partitioning,
    and M^{-1} v matches the oracle's LU_only chain with it."""
    n = 1500
    rp, ci, val = sparse_band(n, 60, 6, 56)
    ref = G.reorder(rp, ci, val, prec="double")
    k = ref["k"]
    P_req = n // (k + 1)
    P = n // (2 * k)
    assert P < P_req, (n, k)
    opts = sap.Options(precValueType=sap.PREC_DOUBLE, factMethod=sap.LU_only, variableBandwidth=int(variable))
    g = sap.Solver(P_req, opts)
    g.setup(sap.CsrMatrix.from_host(rp, ci, val))
    st = g.getStats()
    assert g.numPartitions == st.numPartitions == P and st.bandwidth == k
    if variable:
        var = G.assemble_var(n, k, P, ref["rp"], ref["ci"], ref["val"], prec="double")
        nukf = (2.0 * k * n - (var["ks_row"].sum() + var["ks_col"].sum())) / (2.0 * k * n)
        assert st.nuKf == pytest.approx(nukf, rel=1e-12)
    band = G.csr_to_band(n, k, ref["rp"], ref["ci"], ref["val"])
    cpu = oracle.Solver(P_req, fact="LU_only", prec_double=True)
    assert cpu.setup(band) == 0 and cpu.num_partitions == P
    v = oracle.gen_vector(n, 57)
    w = np.zeros(n)
    w[ref["row_perm"]] = v * ref["row_scale"]
    expect = np.zeros(n)
    expect[ref["col_reord"]] = cpu.apply(w) * ref["col_scale"][ref["col_reord"]]
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    g.precond_apply(dev(v), z)
    np.testing.assert_allclose(host(z), expect, rtol=0, atol=1e-10 * np.abs(expect).max())


def test_csr_testdb_setup(sap):
    """testDB (sap/precond.h:1741-1756): the setup runs DB, reports the
    dominance statistics the host reordering computes and the bandwidth
    after DB, builds no preconditioner (a solve is then Illegal_solve)."""
    rp, ci, val = sparse_band(2000, 40, 5, 70)
    A = sap.CsrMatrix.from_host(rp, ci, val)
    g = sap.Solver(2, sap.Options(testDB=1))
    g.setup(A)
    st = g.getStats()
    r = sap.reorder_csr(rp, ci, val, sap.Options(testDB=1))
    assert (st.diag_dom_ori, st.d_p1_ori, st.diag_dom, st.d_p1) == (r["diag_dom_ori"], r["d_p1_ori"],
                                                                    r["diag_dom"], r["d_p1"])
    assert st.bandwidthDB == r["k_db"] and st.d_p1 > st.d_p1_ori > 0
    b = torch.ones(2000, dtype=torch.float64, device="cuda")
    with pytest.raises(sap.SapError) as ei:
        g.solve(None, b, torch.zeros_like(b))
    assert ei.value.reason == -4


# ------------------------------------------------------------------- update
def _perturb(val, seed):
    return val * (1 + np.random.default_rng(seed).uniform(-0.05, 0.05, val.size))


@pytest.mark.parametrize("prec", ["float", "double"])
@pytest.mark.parametrize("o", [dict(), dict(drop_off=0.01), dict(scale=False)], ids=["db-rcm", "drop", "noscale"])
def test_update_reordered_csr_matches_oracle(sap, oracle, o, prec):
    """update() keeps the setup's permutations and scale factors and only
    refills the band (sap/precond.h:751-775): the oracle's expected band is
    the new value of every entry that setup kept, times its scale factor
    T(r_i c_j), at its reordered place."""
    n, P = 1200, 3
    rp, ci, val = sparse_band(n, 50, 5, 81)
    T = np.float32 if prec == "float" else np.float64
    ref = G.reorder(rp, ci, val, prec=prec, **o)
    k = ref["k"]
    opts = _opts(sap, o, "LU_UL", prec)
    opts.trackReordering = 1
    g = sap.Solver(P, opts)
    A = sap.CsrMatrix.from_host(rp, ci, val)
    g.setup(A)
    val2 = _perturb(val, 82)
    assert g.update(dev(val2))

    scale = o.get("scale", True)
    kept = set(zip(np.repeat(np.arange(n), np.diff(ref["rp"])).tolist(), ref["ci"].tolist()))
    colperm = np.empty(n, dtype=np.int64)
    colperm[ref["col_reord"]] = np.arange(n)
    band = np.zeros((n, 2 * k + 1))
    rows = np.repeat(np.arange(n), np.diff(rp))
    for i, j, a in zip(rows, ci, val2):
        r, c = int(ref["row_perm"][i]), int(colperm[j])
        if (r, c) not in kept:
            continue
        f = T(ref["row_scale_d"][i] * ref["col_scale_d"][j]) if scale else T(1)
        band[r, k + c - r] = float(T(a) * f)
    cpu = oracle.Solver(P, fact="LU_UL", prec_double=(prec == "double"))
    assert cpu.setup(band) == 0

    v = oracle.gen_vector(n, 83)
    w = np.zeros(n)
    w[ref["row_perm"]] = (v.astype(T) * (ref["row_scale"].astype(T) if scale else T(1))).astype(np.float64)
    zr = cpu.apply(w)
    expect = np.zeros(n)
    y = zr.astype(T) * (ref["col_scale"][ref["col_reord"]].astype(T) if scale else T(1))
    expect[ref["col_reord"]] = y.astype(np.float64)
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    g.precond_apply(dev(v), z)
    if prec == "double":
        atol = 1e-10 * np.abs(expect).max()
    else:
        cpu64 = oracle.Solver(P, fact="LU_UL", prec_double=True)
        assert cpu64.setup(band) == 0
        w64 = np.zeros(n)
        w64[ref["row_perm"]] = v * (ref["row_scale"] if scale else 1.0)
        z64 = np.zeros(n)
        z64[ref["col_reord"]] = cpu64.apply(w64) * (ref["col_scale"][ref["col_reord"]] if scale else 1.0)
        atol = max(3e-5 * np.abs(expect).max(), 4 * np.abs(expect - z64).max())
    np.testing.assert_allclose(host(z), expect, rtol=0, atol=atol)
    # the built-in SpMV now multiplies by the new values
    import scipy.sparse as sp
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    b = sp.csr_matrix((val2, ci, rp), shape=(n, n)) @ np.ones(n)
    g.solve(None, dev(b), x)
    assert g.getStats().numSpmv > 0 and g.getStats().timeUpdate > 0


def test_update_banded_and_plain_csr_equal_a_fresh_setup(sap, oracle):
    n, k, P = 4000, 30, 4
    band1 = oracle.gen_banded(n, k, 0.9, 91)
    band2 = oracle.gen_banded(n, k, 1.1, 92)
    opts = sap.Options(factMethod=sap.LU_UL, trackReordering=1)
    g = sap.Solver(P, opts)
    g.setup(sap.BandedMatrix(dev(band1)))
    assert g.update(dev(band2))
    fresh = sap.Solver(P, sap.Options(factMethod=sap.LU_UL))
    fresh.setup(sap.BandedMatrix(dev(band2)))
    np.testing.assert_array_equal(g.factors(), fresh.factors())

    rp, ci, val = csr_of_band(band1)
    _, _, val2 = csr_of_band(band2)
    copts = sap.Options(factMethod=sap.LU_UL, performReorder=0, performDB=0, applyScaling=0,
                        variableBandwidth=0, trackReordering=1)
    gc = sap.Solver(P, copts)
    gc.setup(sap.CsrMatrix.from_host(rp, ci, val))
    assert gc.update(dev(val2))
    np.testing.assert_array_equal(gc.factors(), fresh.factors())
    assert not gc.update(dev(val2[:-1]))           # wrong number of entries: false


def test_update_errors(sap, oracle):
    band = oracle.gen_banded(500, 5, 1.0, 93)
    g = sap.Solver(2, sap.Options(trackReordering=1))
    with pytest.raises(sap.SapError) as ei:
        g.update(dev(band))                         # before setup
    assert ei.value.reason == -3
    g2 = sap.Solver(2, sap.Options())
    g2.setup(sap.BandedMatrix(dev(band)))
    with pytest.raises(sap.SapError) as ei:
        g2.update(dev(band))                        # trackReordering off
    assert ei.value.reason == -3
