"""The general sparse workload (N=500,000, a random band of width 300 with
11 off-diagonal offsets per row, lognormal values with random signs, 20% of
the diagonal scaled by 1e-3, a symmetric random permutation; DB, RCM, 1%
drop-off, SaP-C with 32 partitions) on the CPU:

  * the generator is deterministic and has the stated shape;
  * the shipped host reordering (sap_reorder_csr) is bit-exact with the
    oracle's restatement (oracle/graph_oracle.py) on scaled-down draws of
    the same generator at the full width 300 (permutations, scalings,
    bandwidths, the reordered matrix), in FP32 and FP64;
  * truncated SPIKE on the reordered band is the limit, not the build: the
    oracle's SaP-C (a restatement of the reference's preconditioner and
    BiCGStab(2)) on the reordered, dropped-off band of a 30,000-row draw
    makes no progress with 32 partitions (relative residual 1 after its
    iterations, the same as the GPU solve at full size) and converges with
    two.
"""
import numpy as np
import pytest

from oracle import graph_oracle as G
from saplibrary_amd import workloads as W


@pytest.fixture(scope="module")
def sap():
    import os
    import saplibrary_amd
    if not os.path.exists(saplibrary_amd.LIB_PATH):
        pytest.skip("libsap_amd.so not built")
    saplibrary_amd.lib()
    return saplibrary_amd


def test_generator_shape_and_determinism():
    rp, ci, val = W.general_sparse(50000)
    rp2, ci2, val2 = W.general_sparse(50000)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2) and np.array_equal(val, val2)
    n = 50000
    assert rp.dtype == np.int32 and ci.dtype == np.int32 and rp[0] == 0 and rp[-1] == ci.size
    # ~11.85 entries per row: 11 distinct offsets minus those falling outside
    assert 11.5 < ci.size / n <= 12.0
    # every row has its diagonal, columns sorted within a row
    rows = np.repeat(np.arange(n), np.diff(rp))
    assert np.count_nonzero(rows == ci) == n
    assert np.all((np.diff(ci) > 0) | (np.diff(rows) != 0))
    # 20% of the diagonals shrunk by 1e-3 (lognormal(0,1) * 1e-3 is almost
    # always below 0.05; an unshrunk lognormal value rarely is)
    d = np.abs(val[rows == ci])
    assert 0.18 < np.mean(d < 0.05) < 0.24
    assert 0.47 < np.mean(val < 0) < 0.53


@pytest.mark.parametrize("prec", ["float", "double"])
@pytest.mark.parametrize("n,seed", [(1500, 20), (2500, 23)])
def test_reordering_bit_exact_at_full_width(sap, n, seed, prec):
    rp, ci, val = W.general_sparse(n, seed=seed, weak_seed=seed + 1)
    ref = G.reorder(rp, ci, val, prec=prec, drop_off=0.01)
    opts = sap.Options(dropOffFraction=0.01, precValueType=sap.PREC_FLOAT if prec == "float" else sap.PREC_DOUBLE)
    got = sap.reorder_csr(rp, ci, val, opts)
    for f in ("k", "k_db", "k_reorder"):
        assert got[f] == ref[f], f
    assert got["dropoff"] == ref["dropoff"]
    for f in ("db_perm", "row_perm", "col_reord", "rp", "ci", "row_scale", "col_scale", "val"):
        assert np.array_equal(got[f], ref[f]), f


def test_truncated_spike_limit_on_the_reordered_band(sap, oracle):
    n = 30000
    rp, ci, val = W.general_sparse(n)
    r = sap.reorder_csr(rp, ci, val, sap.Options(dropOffFraction=0.01, precValueType=sap.PREC_DOUBLE))
    band = G.csr_to_band(n, r["k"], r["rp"], r["ci"], r["val"])
    b = oracle.band_spmv(band, np.random.default_rng(5).uniform(-1, 1, n))
    res = {}
    for P in (32, 2):
        S = oracle.Solver(P, rel_tol=1e-10, fact="LU_only", precond="Spike", prec_double=True, refine=3,
                          max_iters=30)
        assert S.setup(band) == 0
        ok, x = S.solve(b)
        rel = np.linalg.norm(b - oracle.band_spmv(band, x)) / np.linalg.norm(b)
        res[P] = (S.num_partitions, ok, rel)
    assert res[32][0] == 32 and not res[32][1] and res[32][2] > 0.99
    assert res[2][0] == 2 and res[2][1] and res[2][2] <= 1e-10
