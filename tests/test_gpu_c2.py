"""The general sparse workload at its full size on the GPU, through the C ABI
(CSR N=500,000, ~12 entries per row, random band of width 300 under a
symmetric permutation, 20% weak diagonals; saplibrary_amd.workloads).

  * The host reordering at full size: DB and RCM permutations are
    permutations, the half-bandwidth after drop-off is the one the solver
    reports, the dropped fraction stays within the 1% asked for.
  * The specified recipe (DB + scaling, RCM, 1% drop-off, SaP-C P=32, FP32
    factors, variable bandwidth) sets up and ends with the honest failure
    code: the truncated SPIKE operator of this matrix is no preconditioner
    at 32 partitions (the oracle behaves the same on the reordered band,
    tests/test_c2_workload.py), so the solve stops with a failure code (its
    refinement limit) and the residual it reached.
  * The configurations that converge (no drop-off, P=4 or 8, FP64 factors)
    reaches a true relative residual <= 1e-10, checked with the library's
    CSR SpMV on the input matrix.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sap():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import saplibrary_amd
    saplibrary_amd.lib()
    return saplibrary_amd


@pytest.fixture(scope="module")
def c2(sap):
    from saplibrary_amd import workloads as W
    n = 500_000
    rp, ci, val = W.general_sparse(n)
    A = sap.CsrMatrix.from_host(rp, ci, val)
    xt = torch.from_numpy(W.uniform_vector(n, 22)).cuda()
    b = torch.empty_like(xt)
    sap.csr_spmv(A, xt, b)
    return rp, ci, val, A, b


def test_c2_reordering_at_full_size(sap, c2):
    rp, ci, val, A, b = c2
    n = rp.size - 1
    print("[c2] host reordering, N=500k, 1% drop-off", flush=True)
    r = sap.reorder_csr(rp, ci, val, sap.Options(dropOffFraction=0.01))
    for key in ("db_perm", "row_perm", "col_reord"):
        assert np.array_equal(np.sort(r[key]), np.arange(n)), key
    assert r["k"] <= r["k_reorder"] < r["k_db"]
    assert 0 < r["dropoff"] <= 0.01
    g = sap.Solver(32, sap.Options(dropOffFraction=0.01))
    g.setup(A)
    st = g.getStats()
    assert (st.bandwidth, st.bandwidthReorder, st.bandwidthDB) == (r["k"], r["k_reorder"], r["k_db"])
    assert st.actualDropOff == pytest.approx(r["dropoff"], rel=1e-6) and st.numPartitions == 32


def test_c2_specified_recipe_fails_honestly(sap, c2):
    rp, ci, val, A, b = c2
    print("[c2] specified recipe: SaP-C P=32 FP32, 1% drop-off", flush=True)
    g = sap.Solver(32, sap.Options(relTol=1e-10, dropOffFraction=0.01, maxNumIterations=30, refine=3))
    g.setup(A)
    x = torch.zeros_like(b)
    ok = g.solve(None, b, x)
    print(f"[c2] specified: code {g.getMonitorCode()} ({g.getMonitorMessage()}), relative residual "
          f"{g.getStats().relResidualNorm:.3e}", flush=True)
    assert not ok and g.getMonitorCode() < 0, (g.getMonitorCode(), g.getMonitorMessage())
    assert np.isfinite(g.getStats().relResidualNorm)


@pytest.mark.parametrize("P", [4, 8])
def test_c2_converging_configuration(sap, c2, P):
    rp, ci, val, A, b = c2
    print(f"[c2] converging configuration: SaP-C P={P} FP64, no drop-off", flush=True)
    g = sap.Solver(P, sap.Options(relTol=1e-10, precValueType=sap.PREC_DOUBLE, refine=4))
    g.setup(A)
    x = torch.zeros_like(b)
    assert g.solve(None, b, x), g.getMonitorMessage()
    r = torch.empty_like(b)
    sap.csr_spmv(A, x, r)
    rel = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
    assert rel <= 1e-10, rel
    assert g.getStats().numPartitions == P
