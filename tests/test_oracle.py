"""Pins the CPU oracle (oracle/sap_oracle.cpp) against closed forms.

Each check is chosen so that a plausible mistake in the restatement (a
transposed spike block, a dropped purification term, a wrong partition
boundary, a missing pivot division, a wrong boosting sign) fails it:
  * generator and SpMV: independent numpy/Python computations + golden fixture
  * FP64 band LU with partial pivoting: scipy.linalg.solve_banded
  * band LU/UL factors: L*D*U (resp. U*D*L) reproduces each partition block
  * SaP-C apply: dense truncated SPIKE built from its definition in numpy
  * SaP-D apply: dense block-diagonal solve; P=1: M = A
  * Krylov + monitor: convergence contract of the reference unit test
    (examples/unit_test/driver_unit_test.cu:87-93), MINRES on SPD systems,
    and the failure codes of sap/monitor.h:291-313.
"""
import os

import numpy as np
import pytest
import scipy.linalg

from tests import spike_dense as SD

HERE = os.path.dirname(os.path.abspath(__file__))

MASK64 = (1 << 64) - 1


def _splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return x, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK64


class _Xoshiro:
    """xoshiro256** as published by Blackman & Vigna (prng.di.unimi.it)."""

    def __init__(self, state):
        self.s = list(state)

    def next(self):
        s = self.s
        result = (_rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result


def _row_stream(seed, row):
    x = seed ^ (((row + 1) * 0xD1B54A32D192ED03) & MASK64)
    st = []
    for _ in range(4):
        x, z = _splitmix64(x)
        st.append(z)
    return _Xoshiro(st)


def test_xoshiro_known_answer():
    # First output of xoshiro256** from state {1,2,3,4} is rotl(2*5,7)*9 = 11520.
    g = _Xoshiro([1, 2, 3, 4])
    assert g.next() == 11520


def test_generator_matches_independent_python(oracle):
    n, k, d, seed = 40, 3, 1.2, 0
    band = oracle.gen_banded(n, k, d, seed)
    ref = np.zeros_like(band)
    for i in range(n):
        g = _row_stream(seed, i)
        s = 0.0
        for dd in range(2 * k + 1):
            if dd == k:
                continue
            j = i + dd - k
            if 0 <= j < n:
                v = (g.next() >> 11) * 2.0 ** -52 - 1.0
                ref[i, dd] = v
                s += abs(v)
        ref[i, k] = d * s
    np.testing.assert_array_equal(band, ref)


def test_generator_golden_fixture(oracle):
    gold = np.load(os.path.join(HERE, "golden", "generator_small.npz"))
    np.testing.assert_array_equal(oracle.gen_banded(40, 3, 1.2, 0), gold["band_n40_k3_d12_s0"])
    np.testing.assert_array_equal(oracle.gen_vector(40, 1), gold["vec_n40_s1"])
    np.testing.assert_array_equal(oracle.gen_banded(300, 12, 0.8, 10), gold["band_n300_k12_d08_s10"])


def test_generator_properties(oracle):
    n, k = 500, 7
    band = oracle.gen_banded(n, k, 0.8, 3)
    off = band.copy()
    off[:, k] = 0
    assert np.all(np.abs(off) <= 1.0)
    np.testing.assert_allclose(band[:, k], 0.8 * np.abs(off).sum(axis=1), rtol=1e-15)
    # out-of-matrix corners are zero
    for i in range(k):
        assert np.all(band[i, :k - i] == 0)
        assert np.all(band[n - 1 - i, k + 1 + i:] == 0)
    assert not np.array_equal(band, oracle.gen_banded(n, k, 0.8, 4))


def test_band_spmv_vs_dense(oracle):
    band = oracle.gen_banded(200, 9, 1.0, 5)
    x = oracle.gen_vector(200, 6)
    A = SD.dense_from_band(band)
    np.testing.assert_allclose(oracle.band_spmv(band, x), A @ x, rtol=1e-13, atol=1e-13)
    gold = np.load(os.path.join(HERE, "golden", "generator_small.npz"))
    np.testing.assert_allclose(oracle.band_spmv(gold["band_n40_k3_d12_s0"], gold["vec_n40_s1"]),
                               gold["spmv_n40"], rtol=0, atol=0)


@pytest.mark.parametrize("n,k,d", [(150, 6, 0.3), (400, 20, 0.8), (60, 1, 0.5)])
def test_band_solve_pp_vs_scipy(oracle, n, k, d):
    band = oracle.gen_banded(n, k, d, 7)
    b = oracle.gen_vector(n, 8)
    ab = np.zeros((2 * k + 1, n))
    for i in range(n):
        for dd in range(2 * k + 1):
            j = i + dd - k
            if 0 <= j < n:
                ab[k + i - j, j] = band[i, dd]
    ref = scipy.linalg.solve_banded((k, k), ab, b)
    np.testing.assert_allclose(oracle.band_solve_pp(band, b), ref, rtol=1e-9, atol=1e-11)


def _lu_check_partition(F, A, s, ln, ul=False):
    blk = F[s:s + ln, s:s + ln]
    D = np.diag(np.diag(blk))
    Lo = np.tril(blk, -1) + np.eye(ln)
    Up = np.triu(blk, 1) + np.eye(ln)
    prod = (Up @ D @ Lo) if ul else (Lo @ D @ Up)
    np.testing.assert_allclose(prod, A[s:s + ln, s:s + ln], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("P,k", [(1, 5), (4, 5), (3, 72)])
def test_block_lu_factors(oracle, P, k):
    n = 480 if k < 70 else 450
    band = oracle.gen_banded(n, k, 0.9, 11)
    S = oracle.Solver(P, precond="Block", prec_double=True)
    assert S.setup(band) == 0
    F = SD.unpack_colband(S.factors(), k)
    A = SD.dense_from_band(band)
    for s, ln in SD.partitions(n, S.num_partitions):
        _lu_check_partition(F, A, s, ln)


def test_lu_ul_factors_last_partition_is_ul(oracle):
    n, k, P = 240, 6, 4
    band = oracle.gen_banded(n, k, 0.8, 12)
    S = oracle.Solver(P, precond="Spike", fact="LU_UL", prec_double=True)
    assert S.setup(band) == 0
    F = SD.unpack_colband(S.factors(), k)
    A = SD.dense_from_band(band)
    parts = SD.partitions(n, P)
    for p, (s, ln) in enumerate(parts):
        _lu_check_partition(F, A, s, ln, ul=(p == P - 1))


@pytest.mark.parametrize("fact", ["LU_UL", "LU_only"])
@pytest.mark.parametrize("n,k,P,d", [(120, 4, 5, 0.8), (333, 7, 6, 1.1), (90, 3, 40, 0.6), (200, 12, 3, 2.0)])
def test_spike_apply_vs_dense_definition(oracle, fact, n, k, P, d):
    band = oracle.gen_banded(n, k, d, 21)
    f = oracle.gen_vector(n, 22)
    S = oracle.Solver(P, precond="Spike", fact=fact, prec_double=True)
    assert S.setup(band) == 0
    Pe = SD.clamp_partitions(n, k, P)
    assert S.num_partitions == Pe
    ref = SD.spike_apply(SD.dense_from_band(band), k, Pe, f)
    np.testing.assert_allclose(S.apply(f), ref, rtol=1e-10, atol=1e-12)


def test_spike_tips_and_reduced_matrix(oracle):
    n, k, P = 160, 5, 4
    band = oracle.gen_banded(n, k, 0.8, 31)
    S = oracle.Solver(P, precond="Spike", fact="LU_UL", prec_double=True)
    S.setup(band)
    R, WV = S.reduced()
    A = SD.dense_from_band(band)
    tips = SD.spike_tips(A, k, SD.partitions(n, P))
    for i, (V, W, _, _) in enumerate(tips):
        # WV stores each K x K tip column by column (RHS-major)
        np.testing.assert_allclose(WV[i, 0].T, V, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(WV[i, 1].T, W, rtol=1e-10, atol=1e-13)
        # R is the LU of [[I, V],[W, I]] (column-major storage): check L*D*U
        Rm = R[i].T
        Lo = np.tril(Rm, -1) + np.eye(2 * k)
        D = np.diag(np.concatenate([np.ones(k), np.diag(Rm)[k:]]))
        Up = np.triu(Rm, 1)
        Up[k:, k:] = np.triu(Rm[k:, k:], 1)
        Up = Up + np.eye(2 * k)
        # rows < k of U hold V undivided (unit pivots), rows >= k are divided
        full = np.block([[np.eye(k), V], [W, np.eye(k)]])
        np.testing.assert_allclose(Lo @ D @ Up, full, rtol=1e-10, atol=1e-12)


def test_spike_fp32_close_to_fp64(oracle):
    n, k, P = 600, 10, 6
    band = oracle.gen_banded(n, k, 0.8, 41)
    f = oracle.gen_vector(n, 42)
    S32 = oracle.Solver(P, precond="Spike", fact="LU_UL")
    S32.setup(band)
    ref = SD.spike_apply(SD.dense_from_band(band), k, P, f)
    np.testing.assert_allclose(S32.apply(f), ref, rtol=2e-4, atol=2e-5 * np.abs(ref).max())


@pytest.mark.parametrize("n,k,P", [(200, 6, 4), (64, 3, 1), (300, 10, 7)])
def test_block_apply_and_p1(oracle, n, k, P):
    band = oracle.gen_banded(n, k, 0.7, 51)
    f = oracle.gen_vector(n, 52)
    A = SD.dense_from_band(band)
    S = oracle.Solver(P, precond="Block", prec_double=True)
    S.setup(band)
    np.testing.assert_allclose(S.apply(f), SD.block_apply(A, k, S.num_partitions, f), rtol=1e-10, atol=1e-12)
    if P == 1:
        np.testing.assert_allclose(S.apply(f), np.linalg.solve(A, f), rtol=1e-10, atol=1e-12)


def test_k0_is_diagonal_divide(oracle):
    n = 50
    band = np.ascontiguousarray(oracle.gen_banded(n, 0, 1.0, 1)[:, :1] + 2.0)
    f = oracle.gen_vector(n, 2)
    S = oracle.Solver(4, precond="Spike", prec_double=True)
    S.setup(band)
    np.testing.assert_allclose(S.apply(f), f / band[:, 0], rtol=1e-15)


def test_partition_clamp(oracle):
    # P <= max(1, N / max(K+1, 2K))  (sap/precond.h:2116-2117)
    band = oracle.gen_banded(100, 10, 1.0, 3)
    S = oracle.Solver(64, precond="Spike")
    S.setup(band)
    assert S.num_partitions == 5


def test_boosting_special_cases(oracle):
    # K >= 70: every pivot boosted; 0 -> +1e-4, -5e-8 -> -1e-4, kept otherwise
    n, k = 300, 70
    band = oracle.gen_banded(n, k, 1.0, 61)
    band[0, k] = 0.0
    S = oracle.Solver(1, precond="Block", prec_double=True)
    S.setup(band)
    assert S.factors()[0, k] == 1e-4
    band[0, k] = -5e-8
    S.setup(band)
    assert S.factors()[0, k] == -1e-4
    band[0, k] = 3e-7
    S.setup(band)
    assert S.factors()[0, k] == 3e-7


def test_zero_pivot_raises_without_safe_factorization(oracle):
    # K < 70, safeFactorization off -> Zero_pivoting (-1); with it -> boosted
    n, k = 100, 5
    band = oracle.gen_banded(n, k, 1.0, 62)
    band[0, k] = 0.0
    S = oracle.Solver(2, precond="Block", prec_double=True)
    assert S.setup(band) == -1
    S2 = oracle.Solver(2, precond="Block", prec_double=True, safe=True)
    assert S2.setup(band) == 0
    assert S2.factors()[0, k] == 1e-4


def test_unit_test_contract_fp64_preconditioner(oracle):
    # examples/unit_test/driver_unit_test.cu:49-94: N=10000, K=20, P=10, d=1.0,
    # LU_UL, relTol 1e-13 -> code 1, "Converged", rel. residual <= 1e-13.
    n, k = 10000, 20
    band = oracle.gen_banded(n, k, 1.0, 70)
    xt = oracle.gen_vector(n, 71)
    b = oracle.band_spmv(band, xt)
    S = oracle.Solver(10, rel_tol=1e-13, fact="LU_UL", prec_double=True)
    assert S.setup(band) == 0
    ok, x = S.solve(b)
    st = S.stats()
    assert ok == 1 and st["code"] == 1 and S.message() == "Converged"
    assert st["rel_residual_norm"] <= 1e-13
    assert np.linalg.norm(b - oracle.band_spmv(band, x)) / np.linalg.norm(b) <= 1e-13


def test_fp32_preconditioner_stagnates_without_refinement(oracle):
    # The right-preconditioned BiCGStab(2) returns x = M^{-1} xx; with FP32
    # factors its rounding bounds the true residual near 1e-7.
    n, k = 10000, 20
    band = oracle.gen_banded(n, k, 1.0, 70)
    b = oracle.band_spmv(band, oracle.gen_vector(n, 71))
    S = oracle.Solver(10, rel_tol=1e-13, fact="LU_UL")
    S.setup(band)
    ok, _ = S.solve(b)
    st = S.stats()
    assert ok == 0 and st["code"] in (-3, -1)
    assert 1e-9 < st["rel_residual_norm"] < 1e-5
    R = oracle.Solver(10, rel_tol=1e-13, fact="LU_UL", refine=6)
    R.setup(band)
    ok, x = R.solve(b)
    assert ok == 1 and R.stats()["code"] == 1 and R.message() == "Converged"
    assert np.linalg.norm(b - oracle.band_spmv(band, x)) / np.linalg.norm(b) <= 1e-13


def test_forward_error_vs_fp64_band_oracle(oracle):
    n, k = 3000, 15
    band = oracle.gen_banded(n, k, 0.8, 80)
    b = oracle.gen_vector(n, 81)
    xref = oracle.band_solve_pp(band, b)
    S = oracle.Solver(8, rel_tol=1e-12, refine=6)
    S.setup(band)
    ok, x = S.solve(b)
    assert ok == 1
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) < 1e-10


def test_monitor_codes(oracle):
    n, k = 2000, 8
    band = oracle.gen_banded(n, k, 0.6, 90)
    b = oracle.gen_vector(n, 91)
    # iteration limit -> -1 (or stagnation -3 before it): never "Converged"
    S = oracle.Solver(4, rel_tol=1e-30, max_iters=0, precond="Block", prec_double=True)
    S.setup(band)
    ok, _ = S.solve(b)
    assert ok == 0 and S.stats()["code"] == -1
    assert S.message() == "Maximum number of iterations was reached"
    # NaN in the right-hand side -> -2
    bn = b.copy()
    bn[5] = np.nan
    S2 = oracle.Solver(4, rel_tol=1e-10, precond="Block", prec_double=True)
    S2.setup(band)
    ok, _ = S2.solve(bn)
    assert ok == 0 and S2.stats()["code"] == -2 and S2.message() == "Residual norm is NaN"
    # iteration counts come in quarters (sap/bicgstab2.h:95,227,256,313)
    S3 = oracle.Solver(4, rel_tol=1e-10, precond="Block", prec_double=True)
    S3.setup(band)
    ok, _ = S3.solve(b)
    assert ok == 1 and (S3.stats()["num_iterations"] * 4) % 1 == 0


def test_minres_spd(oracle):
    n, k = 3000, 10
    band = oracle.gen_banded(n, k, 0.0, 95)
    A = SD.dense_from_band(band)
    A = 0.5 * (A + A.T)
    A[np.diag_indices(n)] = np.abs(A).sum(axis=1) * 1.1 + 0.1
    bandS = np.ascontiguousarray(SD.band_from_dense(A, k))
    b = oracle.gen_vector(n, 96)
    S = oracle.Solver(6, solver="minres", rel_tol=1e-10, precond="Block", prec_double=True, max_iters=200)
    S.setup(bandS)
    ok, x = S.solve(b)
    assert ok == 1 and S.stats()["code"] == 1
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
