"""Block cyclic reduction (useBCR, sap/precond.h:981-1210, :6485-7076,
sap/segmented_matrix.h) on the GPU, through the C ABI.

BCR factors every partition by cyclic reduction over blocks of at least 2k
rows instead of one sequential banded LU; like the reference's BCR setup it
builds the block-diagonal preconditioner, so M^{-1} v is A_i^{-1} v_i on
every partition.  Pinned against the oracle's Block preconditioner (banded
LU of each partition):

  * FP64: within 1e-10 of max |z| (the two eliminations differ only in
    rounding on these diagonally dominant matrices), over block counts that
    are powers of two and not, one block per partition (plain LU), P = 1 and
    widths from 3 to 120;
  * FP32: normwise within 3e-5 of the oracle's FP64 result;
  * the reordered CSR path (DB, RCM, constant band) against the oracle's
    Block chain;
  * solves to 1e-10 with refinement; several right-hand sides in one batched
    solve equal to the one-column solves bit for bit (no refinement).
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


CASES = [
    # n, k, P, dominance
    (20000, 40, 4, 1.0),    # 5000-row partitions: 62 blocks each (6 levels)
    (6000, 30, 2, 1.2),     # 50 blocks
    (4096, 16, 1, 1.0),     # P = 1: 128 blocks (a power of two)
    (3000, 3, 3, 1.5),      # narrow band, 166 blocks
    (1000, 200, 4, 1.0),    # partitions shorter than 4k: one block each (plain LU)
    (30000, 120, 8, 1.1),   # 15 blocks of >= 240 rows
]


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("n,k,P,d", CASES)
def test_bcr_apply_matches_oracle_block(sap, oracle, n, k, P, d, prec):
    print(f"[bcr] n={n} k={k} P={P} d={d} {prec}", flush=True)
    band = oracle.gen_banded(n, k, d, 500 + k)
    v = oracle.gen_vector(n, 501)
    cpu = oracle.Solver(P, precond="Block", prec_double=True)
    assert cpu.setup(band) == 0
    z64 = cpu.apply(v)
    g = sap.Solver(P, sap.Options(useBCR=1, precondType=sap.Block,
                                  precValueType=sap.PREC_DOUBLE if prec == "double" else sap.PREC_FLOAT))
    g.setup(sap.BandedMatrix(dev(band)))
    assert g.numPartitions == cpu.num_partitions
    if n // g.numPartitions >= 4 * k:
        # several blocks per partition: the reduction updated block corners,
        # so the stored factors are not the plain banded LU's
        plain = sap.Solver(P, sap.Options(precondType=sap.Block,
                                          precValueType=sap.PREC_DOUBLE if prec == "double" else sap.PREC_FLOAT))
        plain.setup(sap.BandedMatrix(dev(band)))
        assert not np.array_equal(g.factors(), plain.factors())
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    g.precond_apply(dev(v), z)
    zg = host(z)
    if prec == "double":
        np.testing.assert_allclose(zg, z64, rtol=0, atol=1e-10 * np.abs(z64).max())
    else:
        assert np.linalg.norm(zg - z64) / np.linalg.norm(z64) <= 3e-5
    # a Spike request with BCR builds the same block-diagonal preconditioner
    gs = sap.Solver(P, sap.Options(useBCR=1, precondType=sap.Spike,
                                   precValueType=sap.PREC_DOUBLE if prec == "double" else sap.PREC_FLOAT))
    gs.setup(sap.BandedMatrix(dev(band)))
    z2 = torch.empty_like(z)
    gs.precond_apply(dev(v), z2)
    torch.cuda.synchronize()
    assert torch.equal(z, z2)


def test_bcr_solve_and_batched_bitwise(sap, oracle):
    n, k, P = 40000, 60, 4
    band = oracle.gen_banded(n, k, 1.0, 510)
    A = sap.BandedMatrix(dev(band))
    b = oracle.band_spmv(band, oracle.gen_vector(n, 511))
    g = sap.Solver(P, sap.Options(useBCR=1, precondType=sap.Block, relTol=1e-10, refine=8))
    g.setup(A)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    assert g.solve(None, dev(b), x), g.getMonitorMessage()
    assert np.linalg.norm(b - oracle.band_spmv(band, host(x))) / np.linalg.norm(b) <= 1e-10
    # three right-hand sides in one batched solve, no refinement: each column
    # as the one-column solve, bit for bit
    B = np.stack([oracle.gen_vector(n, 512 + c) for c in range(3)], axis=1)
    o = dict(useBCR=1, precondType=sap.Block, relTol=1e-6)
    g1 = sap.Solver(P, sap.Options(**o))
    g1.setup(A)
    cols = []
    for c in range(3):
        xc = torch.zeros(n, dtype=torch.float64, device="cuda")
        g1.solve(None, dev(B[:, c]), xc)
        cols.append(host(xc))
    gb = sap.Solver(P, sap.Options(batch=1, **o))
    gb.setup(A)
    X = torch.zeros((n, 3), dtype=torch.float64, device="cuda")
    gb.solve(None, dev(B), X)
    np.testing.assert_array_equal(host(X), np.stack(cols, axis=1))


def test_bcr_reordered_csr_matches_oracle_block(sap, oracle):
    n, P = 3000, 3
    rp, ci, val = sparse_band(n, 40, 4, 520, weak_diag=0.0)
    ref = G.reorder(rp, ci, val, prec="double")
    k = ref["k"]
    band = G.csr_to_band(n, k, ref["rp"], ref["ci"], ref["val"])
    cpu = oracle.Solver(P, precond="Block", fact="LU_only", prec_double=True)
    assert cpu.setup(band) == 0
    g = sap.Solver(P, sap.Options(useBCR=1, precondType=sap.Block, precValueType=sap.PREC_DOUBLE))
    g.setup(sap.CsrMatrix.from_host(rp, ci, val))
    assert g.numPartitions == cpu.num_partitions and g.getStats().bandwidth == k
    v = oracle.gen_vector(n, 521)
    w = np.zeros(n)
    w[ref["row_perm"]] = v * ref["row_scale"]
    expect = np.zeros(n)
    expect[ref["col_reord"]] = cpu.apply(w) * ref["col_scale"][ref["col_reord"]]
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    g.precond_apply(dev(v), z)
    np.testing.assert_allclose(host(z), expect, rtol=0, atol=1e-9 * np.abs(expect).max())


def test_bcr_update_equals_fresh_setup(sap, oracle):
    """update() with block cyclic reduction reruns the reduction on the new
    values (same blocks, levels and kernels), so it equals a fresh setup of
    those values bit for bit."""
    n, k, P = 12000, 24, 3
    band1 = oracle.gen_banded(n, k, 1.0, 530)
    band2 = oracle.gen_banded(n, k, 1.1, 531)
    o = dict(useBCR=1, precondType=sap.Block, trackReordering=1)
    g = sap.Solver(P, sap.Options(**o))
    g.setup(sap.BandedMatrix(dev(band1)))
    assert g.update(dev(band2))
    fresh = sap.Solver(P, sap.Options(**o))
    fresh.setup(sap.BandedMatrix(dev(band2)))
    np.testing.assert_array_equal(g.factors(), fresh.factors())
    v = dev(oracle.gen_vector(n, 532))
    z1 = torch.empty_like(v)
    z2 = torch.empty_like(v)
    g.precond_apply(v, z1)
    fresh.precond_apply(v, z2)
    torch.cuda.synchronize()
    assert torch.equal(z1, z2)
