"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Floating-point tolerances (stated per test): FP64 paths within 1e-9
relative (different but fixed summation orders); FP32 preconditioner
quantities within a few hundred FP32 ulps of the factor magnitudes.  Integer
and generator outputs are bit-exact.
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


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make_pair(sap, oracle, P, fact="LU_UL", precond="Spike", prec_double=False, safe=False, **kw):
    opts = sap.Options(factMethod=sap.LU_UL if fact == "LU_UL" else sap.LU_only,
                       precondType={"Spike": sap.Spike, "Block": sap.Block, "None": sap.NoPrecond}[precond],
                       precValueType=sap.PREC_DOUBLE if prec_double else sap.PREC_FLOAT,
                       safeFactorization=int(safe), **kw)
    ref_kw = {}
    if "relTol" in kw:
        ref_kw["rel_tol"] = kw["relTol"]
    if "refine" in kw:
        ref_kw["refine"] = kw["refine"]
    if "maxNumIterations" in kw:
        ref_kw["max_iters"] = kw["maxNumIterations"]
    if kw.get("solverType") == sap.MINRES:
        ref_kw["solver"] = "minres"
    gpu = sap.Solver(P, opts)
    cpu = oracle.Solver(P, fact=fact, precond=precond, prec_double=prec_double, safe=safe, **ref_kw)
    return gpu, cpu


# --------------------------------------------------------------- generators
def test_generator_bit_exact(sap, oracle):
    for (n, k, d, seed) in [(40, 3, 1.2, 0), (1000, 50, 0.8, 10), (777, 0, 2.0, 5)]:
        A = sap.BandedMatrix.generate(n, k, d, seed)
        np.testing.assert_array_equal(host(A.band), oracle.gen_banded(n, k, d, seed))
    np.testing.assert_array_equal(host(sap.gen_vector(5000, 11)), oracle.gen_vector(5000, 11))


@pytest.mark.parametrize("n,k", [(1000, 3), (5000, 50), (3000, 200), (10, 4)])
def test_band_spmv(sap, oracle, n, k):
    band = oracle.gen_banded(n, k, 0.9, 3)
    x = oracle.gen_vector(n, 4)
    A = sap.BandedMatrix(dev(band))
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    sap.band_spmv(A, dev(x), y)
    ref = oracle.band_spmv(band, x)
    np.testing.assert_allclose(host(y), ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


# ----------------------------------------------------------------- factors
FACTOR_CASES = [
    # n, k, P, fact, precond
    (2000, 5, 4, "LU_UL", "Spike"),
    (3000, 40, 6, "LU_UL", "Spike"),
    (6000, 80, 5, "LU_UL", "Spike"),
    (3000, 12, 5, "LU_only", "Spike"),
    (4000, 30, 3, "LU_UL", "Block"),
    (900, 7, 1, "LU_UL", "Spike"),
]


@pytest.mark.parametrize("n,k,P,fact,precond", FACTOR_CASES)
@pytest.mark.parametrize("prec_double", [True, False])
def test_factors_match_oracle(sap, oracle, n, k, P, fact, precond, prec_double):
    band = oracle.gen_banded(n, k, 0.8, 21)
    gpu, cpu = make_pair(sap, oracle, P, fact, precond, prec_double)
    A = sap.BandedMatrix(dev(band))
    gpu.setup(A)
    assert cpu.setup(band) == 0
    assert gpu.numPartitions == cpu.num_partitions
    Fg, Fc = gpu.factors(), cpu.factors()
    scale = np.abs(Fc).max()
    tol = 1e-10 if prec_double else 2e-4
    np.testing.assert_allclose(Fg, Fc, rtol=0, atol=tol * scale)
    if precond == "Spike" and cpu.num_partitions > 1:
        wv, s = gpu.spikes()
        R, WVc = cpu.reduced()
        np.testing.assert_allclose(wv, WVc, rtol=0, atol=tol * max(1.0, np.abs(WVc).max()))
        # the oracle's R_i is 2k x 2k column-major; its trailing block holds S's factors
        for i in range(cpu.num_partitions - 1):
            Sc = R[i].T[k:, k:]
            np.testing.assert_allclose(s[i].T, Sc, rtol=0, atol=tol * max(1.0, np.abs(Sc).max()))


# Bands wider than the register-resident factorization takes: the blocked
# matrix-core kernel with two panel rows per thread (k + 32 > 256), -L_below
# staged in LDS (FP32, and FP64 up to k = 256) or read back from the band
# (FP64 from k = 272 on, where the LDS holds only the U block).
WIDE_FACTOR_CASES = [(4000, 230, 3, "LU_UL", "Block"), (3000, 300, 2, "LU_only", "Block"),
                     (2400, 460, 2, "LU_UL", "Block")]


@pytest.mark.parametrize("n,k,P,fact,precond", WIDE_FACTOR_CASES)
@pytest.mark.parametrize("prec_double", [True, False])
def test_wide_band_factors_match_oracle(sap, oracle, n, k, P, fact, precond, prec_double):
    band = oracle.gen_banded(n, k, 0.8, 25)
    gpu, cpu = make_pair(sap, oracle, P, fact, precond, prec_double)
    gpu.setup(sap.BandedMatrix(dev(band)))
    assert cpu.setup(band) == 0
    assert gpu.numPartitions == cpu.num_partitions == P
    Fg, Fc = gpu.factors(), cpu.factors()
    tol = 1e-10 if prec_double else 2e-4
    np.testing.assert_allclose(Fg, Fc, rtol=0, atol=tol * np.abs(Fc).max())


@pytest.mark.parametrize("n,k,P,fact,precond", FACTOR_CASES)
@pytest.mark.parametrize("prec_double", [True, False])
def test_actual_nnz_counts_the_factored_band(sap, oracle, n, k, P, fact, precond, prec_double):
    """Stats.actual_nnz is (2k+1) n minus the zeros of the factored band
    (sap/precond.h:1352, thrust::count over m_B), the band factors() returns
    (the UL-factored last partition under LU_UL, as in m_B)."""
    band = oracle.gen_banded(n, k, 0.8, 23)
    gpu, _ = make_pair(sap, oracle, P, fact, precond, prec_double)
    gpu.setup(sap.BandedMatrix(dev(band)))
    F = gpu.factors()
    assert gpu.getStats().actual_nnz == F.size - int(np.count_nonzero(F == 0.0))
    # out-of-matrix corners are exact zeros, so the count is below the full band
    assert gpu.getStats().actual_nnz <= F.size - k * (k + 1)


# ------------------------------------------------------------------- apply
APPLY_CASES = FACTOR_CASES + [(20000, 100, 16, "LU_UL", "Spike"), (5000, 0, 4, "LU_UL", "Spike"),
                              (3000, 10, 3, "LU_UL", "None"),
                              # k above the pipelined-sweep limit (224 FP32, 208 FP64) and between them
                              (5000, 300, 4, "LU_UL", "Spike"), (8000, 210, 8, "LU_UL", "Spike"),
                              (7000, 230, 3, "LU_only", "Spike"),
                              # wide FP64 bands (the staging buffers no longer fit the LDS: plain sweeps)
                              (9000, 350, 4, "LU_only", "Block"), (7000, 380, 3, "LU_UL", "Spike")]


@pytest.mark.parametrize("n,k,P,fact,precond", APPLY_CASES)
@pytest.mark.parametrize("prec_double", [True, False])
def test_precond_apply_matches_oracle(sap, oracle, n, k, P, fact, precond, prec_double):
    band = oracle.gen_banded(n, k, 0.8, 31)
    if k == 0:
        band = np.ascontiguousarray(band + 1.5)
    v = oracle.gen_vector(n, 32)
    gpu, cpu = make_pair(sap, oracle, P, fact, precond, prec_double)
    gpu.setup(sap.BandedMatrix(dev(band)))
    cpu.setup(band)
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    gpu.precond_apply(dev(v), z)
    ref = cpu.apply(v)
    tol = 1e-10 if prec_double else 3e-5
    np.testing.assert_allclose(host(z), ref, rtol=0, atol=tol * np.abs(ref).max())


# ------------------------------------------------------------------- solve
SOLVE_CASES = [
    # n, k, P, d, fact, precond, prec_double, refine, tol
    (20000, 20, 10, 1.0, "LU_UL", "Spike", True, 0, 1e-13),    # unit-test contract, FP64 preconditioner
    (20000, 20, 10, 1.0, "LU_UL", "Spike", False, 6, 1e-13),   # FP32 preconditioner + refinement
    (50000, 50, 16, 1.2, "LU_UL", "Block", False, 6, 1e-10),   # SaP-D, FP32
    (30000, 120, 8, 0.8, "LU_UL", "Spike", False, 6, 1e-10),   # SaP-C, K >= 70 (blocked, boosted)
    (20000, 15, 8, 0.8, "LU_only", "Spike", False, 6, 1e-10),
    (20000, 30, 8, 0.8, "LU_UL", "Block", True, 0, 1e-10),     # plain FP64 SaP-D
]


@pytest.mark.parametrize("n,k,P,d,fact,precond,prec_double,refine,tol", SOLVE_CASES)
def test_solve_matches_oracle(sap, oracle, n, k, P, d, fact, precond, prec_double, refine, tol):
    band = oracle.gen_banded(n, k, d, 41)
    xt = oracle.gen_vector(n, 42)
    b = oracle.band_spmv(band, xt)
    gpu, cpu = make_pair(sap, oracle, P, fact, precond, prec_double, relTol=tol, refine=refine)
    A = sap.BandedMatrix(dev(band))
    gpu.setup(A)
    cpu.setup(band)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    ok = gpu.solve(sap.MVBanded(A), dev(b), x)
    okc, xc = cpu.solve(b)
    st, stc = gpu.getStats(), cpu.stats()
    assert ok and okc
    assert gpu.getMonitorCode() == 1 and gpu.getMonitorMessage() == "Converged"
    xg = host(x)
    rel = np.linalg.norm(b - oracle.band_spmv(band, xg)) / np.linalg.norm(b)
    assert rel <= tol
    assert st.relResidualNorm <= tol
    # same iteration count as the oracle within one quarter-step of BiCGStab(2)
    assert abs(st.numIterations - stc["num_iterations"]) <= 0.5
    assert st.numRefine == stc["num_refine"]
    np.testing.assert_allclose(xg, xt, rtol=0, atol=1e-6 * np.abs(xt).max())


def test_forward_error_vs_fp64_band_oracle(sap, oracle):
    n, k = 6000, 25
    band = oracle.gen_banded(n, k, 0.8, 51)
    b = oracle.gen_vector(n, 52)
    xref = oracle.band_solve_pp(band, b)
    gpu, _ = make_pair(sap, oracle, 8, relTol=1e-12, refine=6)
    gpu.setup(sap.BandedMatrix(dev(band)))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert gpu.solve(None, dev(b), x)
    assert np.linalg.norm(host(x) - xref) / np.linalg.norm(xref) < 1e-10


def test_user_spmv_functor(sap, oracle):
    n, k = 8000, 10
    band = oracle.gen_banded(n, k, 1.0, 61)
    b = oracle.gen_vector(n, 62)
    A = sap.BandedMatrix(dev(band))
    calls = []

    def spmv(v, Av):
        calls.append(1)
        sap.band_spmv(A, v, Av)

    gpu, _ = make_pair(sap, oracle, 4, relTol=1e-10, refine=6)
    gpu.setup(A)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert gpu.solve(spmv, dev(b), x)
    assert len(calls) == gpu.getStats().numSpmv > 0
    rel = np.linalg.norm(b - oracle.band_spmv(band, host(x))) / np.linalg.norm(b)
    assert rel <= 1e-10


# ----------------------------------------------------- error behaviour
def test_boosting_and_zero_pivot(sap, oracle):
    n, k = 600, 72
    band = oracle.gen_banded(n, k, 1.0, 71)
    band[0, k] = 0.0
    gpu, _ = make_pair(sap, oracle, 1, precond="Block", prec_double=True)
    gpu.setup(sap.BandedMatrix(dev(band)))
    assert gpu.factors()[0, k] == 1e-4             # K >= 70: boosted, never an error
    band2 = oracle.gen_banded(200, 5, 1.0, 72)
    band2[3, 5] = 0.0
    band2[3, :] = 0.0                               # a zero row -> zero pivot (K < 70, not safe)
    band2 = np.ascontiguousarray(band2)
    g2, _ = make_pair(sap, oracle, 2, precond="Block", prec_double=True)
    with pytest.raises(sap.SapError) as ei:
        g2.setup(sap.BandedMatrix(dev(band2)))
    assert ei.value.reason == -1
    g3, _ = make_pair(sap, oracle, 2, precond="Block", prec_double=True, safe=True)
    g3.setup(sap.BandedMatrix(dev(band2)))


def test_illegal_solve_and_monitor_codes(sap, oracle):
    n, k = 3000, 8
    band = oracle.gen_banded(n, k, 0.6, 81)
    b = oracle.gen_vector(n, 82)
    g, _ = make_pair(sap, oracle, 4)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    with pytest.raises(sap.SapError) as ei:
        g.solve(None, dev(b), x)
    assert ei.value.reason == -4                    # Illegal_solve before setup
    A = sap.BandedMatrix(dev(band))
    g1, _ = make_pair(sap, oracle, 4, precond="Block", prec_double=True, relTol=1e-30, maxNumIterations=0)
    g1.setup(A)
    assert not g1.solve(None, dev(b), x)
    assert g1.getMonitorCode() == -1
    assert g1.getMonitorMessage() == "Maximum number of iterations was reached"
    bn = b.copy()
    bn[7] = np.nan
    g2, _ = make_pair(sap, oracle, 4, precond="Block", prec_double=True, relTol=1e-10)
    g2.setup(A)
    x.zero_()
    assert not g2.solve(None, dev(bn), x)
    assert g2.getMonitorCode() == -2 and g2.getMonitorMessage() == "Residual norm is NaN"


def test_minres_spd(sap, oracle):
    from tests import spike_dense as SD
    n, k = 3000, 10
    band = oracle.gen_banded(n, k, 0.0, 95)
    Ad = SD.dense_from_band(band)
    Ad = 0.5 * (Ad + Ad.T)
    Ad[np.diag_indices(n)] = np.abs(Ad).sum(axis=1) * 1.1 + 0.1
    bandS = SD.band_from_dense(Ad, k)
    b = oracle.gen_vector(n, 96)
    g, c = make_pair(sap, oracle, 6, precond="Block", prec_double=True, relTol=1e-10,
                     solverType=sap.MINRES, maxNumIterations=200)
    g.setup(sap.BandedMatrix(dev(bandS)))
    c.setup(np.ascontiguousarray(bandS))
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert g.solve(None, dev(b), x)
    okc, _ = c.solve(b)
    assert okc
    assert np.linalg.norm(b - Ad @ host(x)) / np.linalg.norm(b) <= 1e-10
    assert abs(g.getStats().numIterations - c.stats()["num_iterations"]) <= 1


def test_large_c0_shape_properties(sap, oracle):
    # c0 shape (N=200k, K=50, P=16, SaP-D, FP32, tol 1e-10): checked through
    # size-independent properties (true residual, forward error vs x_true).
    n, k = 200000, 50
    A = sap.BandedMatrix.generate(n, k, 1.2, 0)
    xt = sap.gen_vector(n, 1)
    b = torch.empty_like(xt)
    sap.band_spmv(A, xt, b)
    g = sap.Solver(16, sap.Options(precondType=sap.Block, relTol=1e-10, refine=6))
    g.setup(A)
    assert g.numPartitions == 16
    x = torch.zeros_like(b)
    assert g.solve(None, b, x)
    r = torch.empty_like(b)
    sap.band_spmv(A, x, r)
    rel = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
    assert rel <= 1e-10
    assert float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt)) < 1e-9
