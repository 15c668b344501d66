"""The headline banded workload (N=1M, K=200, diagonal 0.8x the off-diagonal
abs-sum, SaP-C LU_UL, P=64, FP32 factors) through the C ABI.

  * At an oracle-sized N (64 partitions of 420 rows, the same K, P, d and
    options) the factors, spike tips, Schur factors and M^{-1} v match the
    CPU oracle: FP64 within 1e-10 and FP32 within 2e-4 (factors, tips, Schur
    blocks) / 3e-5 (M^{-1} v) of the largest magnitude.  The oracle's own
    FP32-vs-FP64 difference at this size is ~1.3e-6, so the FP32 bounds are
    loose by design (summation orders differ).
  * Every width of the register-resident factorization's window (16-row
    tiles KT = (K+47)/16 from 3 to 16 in FP32, 3 to 9 in FP64) is compared
    with the oracle, as are partitions shorter than one window (n_p = 2K).
  * At the full size the solve is checked through properties that do not
    depend on N: the FP64 true residual <= 1e-10, the forward error against
    x_true, monitor code 1.
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


def pair(sap, oracle, P, fact, precond, prec_double):
    opts = sap.Options(factMethod=sap.LU_UL if fact == "LU_UL" else sap.LU_only,
                       precondType={"Spike": sap.Spike, "Block": sap.Block}[precond],
                       precValueType=sap.PREC_DOUBLE if prec_double else sap.PREC_FLOAT)
    return sap.Solver(P, opts), oracle.Solver(P, fact=fact, precond=precond, prec_double=prec_double)


def compare_setup(sap, oracle, n, k, P, d, fact, precond, prec_double, seed):
    band = oracle.gen_banded(n, k, d, seed)
    v = oracle.gen_vector(n, seed + 1)
    gpu, cpu = pair(sap, oracle, P, fact, precond, prec_double)
    print(f"[c1 parity] n={n} k={k} P={P} d={d} {fact} {precond} fp64={prec_double}", flush=True)
    gpu.setup(sap.BandedMatrix(dev(band)))
    assert cpu.setup(band) == 0
    assert gpu.numPartitions == cpu.num_partitions
    tol = 1e-10 if prec_double else 2e-4
    Fg, Fc = gpu.factors(), cpu.factors()
    np.testing.assert_allclose(Fg, Fc, rtol=0, atol=tol * np.abs(Fc).max())
    if precond == "Spike" and cpu.num_partitions > 1:
        wv, s = gpu.spikes()
        R, WVc = cpu.reduced()
        np.testing.assert_allclose(wv, WVc, rtol=0, atol=tol * max(1.0, np.abs(WVc).max()))
        for i in range(cpu.num_partitions - 1):
            Sc = R[i].T[k:, k:]
            np.testing.assert_allclose(s[i].T, Sc, rtol=0, atol=tol * max(1.0, np.abs(Sc).max()))
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    gpu.precond_apply(dev(v), z)
    ref = cpu.apply(v)
    atol = (1e-10 if prec_double else 3e-5) * np.abs(ref).max()
    np.testing.assert_allclose(host(z), ref, rtol=0, atol=atol)


@pytest.mark.parametrize("prec_double", [False, True])
def test_c1_shape_at_oracle_size(sap, oracle, prec_double):
    # K=200, P=64, d=0.8, LU_UL, SaP-C: 64 partitions of 420 rows (>= 2K)
    compare_setup(sap, oracle, 64 * 420, 200, 64, 0.8, "LU_UL", "Spike", prec_double, 10)


@pytest.mark.parametrize("prec_double", [False, True])
def test_c1_shape_lu_only_at_oracle_size(sap, oracle, prec_double):
    compare_setup(sap, oracle, 16 * 430, 200, 16, 0.8, "LU_only", "Spike", prec_double, 12)


# one case per window width KT = (k + 47) // 16 not covered elsewhere, and
# partitions shorter than the window (n_p = 2k, small k)
WINDOW_CASES = [
    # n, k, P, fact, precond
    (1200, 88, 4, "LU_UL", "Spike"),      # KT 8
    (1600, 136, 4, "LU_UL", "Spike"),     # KT 11
    (1500, 150, 4, "LU_only", "Spike"),   # KT 12
    (2100, 170, 4, "LU_UL", "Spike"),     # KT 13
    (2000, 185, 4, "LU_UL", "Block"),     # KT 14
    (2400, 200, 5, "LU_UL", "Spike"),     # KT 15 (the headline K)
    (3000, 224, 6, "LU_UL", "Spike"),     # KT 16, the widest resident window
    (200, 10, 10, "LU_UL", "Spike"),      # n_p = 2k = 20 < 16 KT
    (260, 13, 10, "LU_only", "Spike"),    # n_p = 26
    (330, 33, 5, "LU_UL", "Spike"),       # n_p = 66 = 2k, KT 5
]


@pytest.mark.parametrize("n,k,P,fact,precond", WINDOW_CASES)
@pytest.mark.parametrize("prec_double", [False, True])
def test_every_window_width_matches_oracle(sap, oracle, n, k, P, fact, precond, prec_double):
    compare_setup(sap, oracle, n, k, P, 0.8, fact, precond, prec_double, 100 + k)


def test_c1_full_size_properties(sap):
    """N=1M, K=200, d=0.8, SaP-C LU_UL, P=64, FP32 factors, FP64 BiCGStab(2)
    with FP64 refinement, relative tolerance 1e-10."""
    n, k, P = 1_000_000, 200, 64
    print(f"[c1 full] n={n} k={k} P={P}", flush=True)
    A = sap.BandedMatrix.generate(n, k, 0.8, 10)
    xt = sap.gen_vector(n, 11)
    b = torch.empty_like(xt)
    sap.band_spmv(A, xt, b)
    g = sap.Solver(P, sap.Options(precondType=sap.Spike, factMethod=sap.LU_UL, relTol=1e-10, refine=8))
    g.setup(A)
    assert g.numPartitions == P
    x = torch.zeros_like(b)
    assert g.solve(None, b, x), g.getMonitorMessage()
    assert g.getMonitorCode() == 1 and g.getMonitorMessage() == "Converged"
    r = torch.empty_like(b)
    sap.band_spmv(A, x, r)
    rel = float(torch.linalg.norm(b - r) / torch.linalg.norm(b))
    assert rel <= 1e-10
    assert g.getStats().relResidualNorm <= 1e-10
    ferr = float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt))
    assert ferr < 1e-8, ferr
    del A, b, x, r, xt, g
    torch.cuda.empty_cache()
