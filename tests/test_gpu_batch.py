"""The batched solve (several right-hand sides of one system together: one
launch per kernel family for all columns, the BiCGStab(2) scalars and the
monitor on the device, each iteration replayed from a captured HIP graph),
through the C ABI.

  * Without refinement every column is bit-identical to solving that
    column alone with the one-column path (same iterates, same final
    M^{-1} xx, same monitor decisions), with and without the graph, on the
    banded SaP-C (LU_UL and LU_only) and SaP-D paths in FP32 and FP64, on a
    reordered CSR setup, and on the non-converged endings (maximum
    iterations, the best-iterate choice) and stagnation.
  * With FP64 refinement every column reaches the tolerance, and the
    solution agrees with the one-column path to the tolerance (the
    correction solves start from b - A x formed afresh, the one-column path
    from the previous correction's own residual).
  * Aggregate statistics: the largest iteration count over the columns.
"""
import numpy as np
import pytest

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


def gen_band(oracle, n, k, d, seed):
    """The banded generator; for k = 0 (no off-diagonals, so the generator's
    diagonal would be 0) a diagonal in [1, 3)."""
    if k == 0:
        return (oracle.gen_vector(n, seed) + 2.0).reshape(n, 1)
    return oracle.gen_banded(n, k, d, seed)


def one_by_one(sap, A, opts, P, B, X0):
    """Each column solved alone (nrhs = 1): solutions, iteration counts, codes."""
    xs, its, codes = [], [], []
    g = sap.Solver(P, sap.Options(**opts))
    g.setup(A)
    for c in range(B.shape[1]):
        x = dev(X0[:, c])
        g.solve(None, dev(B[:, c]), x)
        xs.append(host(x))
        its.append(g.getStats().numIterations)
        codes.append(g.getMonitorCode())
    return np.stack(xs, axis=1), its, codes


def batched(sap, A, opts, P, B, X0, graph=1):
    g = sap.Solver(P, sap.Options(batch=1, graph=graph, **opts))
    g.setup(A)
    X = dev(X0)
    ok = g.solve(None, dev(B), X)
    return host(X), g.getStats().numIterations, g.getMonitorCode(), ok


BANDED = [
    # n, k, P, dominance, options
    (6000, 30, 4, 0.8, dict(factMethod=0, precondType=0)),                    # SaP-C LU_UL FP32
    (5000, 24, 5, 0.9, dict(factMethod=1, precondType=0)),                    # SaP-C LU_only FP32
    (7000, 12, 4, 1.2, dict(factMethod=0, precondType=1)),                    # SaP-D FP32
    (4000, 40, 3, 0.8, dict(factMethod=0, precondType=0, precValueType=1)),   # SaP-C FP64
    (3000, 0, 2, 2.0, dict(precondType=1)),                                   # K = 0: diagonal divide
]


@pytest.mark.parametrize("n,k,P,d,opts", BANDED)
@pytest.mark.parametrize("graph", [1, 0])
def test_batched_equals_one_by_one_bitwise(sap, oracle, n, k, P, d, opts, graph):
    print(f"[batch] n={n} k={k} P={P} d={d} {opts} graph={graph}", flush=True)
    band = gen_band(oracle, n, k, d, 300 + k)
    nr = 5
    B = np.stack([oracle.gen_vector(n, 310 + c) for c in range(nr)], axis=1)
    X0 = np.zeros_like(B)
    X0[:, 1] = oracle.gen_vector(n, 399) * 0.1     # one column with a nonzero guess
    A = sap.BandedMatrix(dev(band))
    o = dict(relTol=1e-6 if opts.get("precValueType", 0) == 0 else 1e-10, **opts)
    Xs, its, codes = one_by_one(sap, A, o, P, B, X0)
    Xb, itb, codeb, ok = batched(sap, A, o, P, B, X0, graph)
    np.testing.assert_array_equal(Xb, Xs)
    assert itb == max(its)
    # the code of the first column that did not converge, else 1
    assert codeb == next((c for c in codes if c <= 0), codes[-1])
    assert ok == all(c == 1 for c in codes)


@pytest.mark.parametrize("graph", [1, 0])
def test_batched_reordered_csr_bitwise(sap, graph):
    n = 3000
    rp, ci, val = sparse_band(n, 60, 6, 71)
    A = sap.CsrMatrix.from_host(rp, ci, val)
    rng = np.random.default_rng(72)
    B = rng.uniform(-1, 1, (n, 4))
    X0 = np.zeros_like(B)
    o = dict(relTol=1e-6, maxNumIterations=60)
    Xs, its, codes = one_by_one(sap, A, o, 2, B, X0)
    Xb, itb, codeb, _ = batched(sap, A, o, 2, B, X0, graph)
    np.testing.assert_array_equal(Xb, Xs)
    assert itb == max(its)


def test_batched_non_converged_endings_bitwise(sap, oracle):
    """Maximum iterations (-1) with the best-iterate choice, and stagnation
    (-3) of an FP32 preconditioner asked for 1e-14: the same codes, counts
    and returned iterates as the one-column path."""
    n, k, P = 5000, 20, 8
    band = oracle.gen_banded(n, k, 0.55, 321)
    A = sap.BandedMatrix(dev(band))
    B = np.stack([oracle.gen_vector(n, 322 + c) for c in range(3)], axis=1)
    X0 = np.zeros_like(B)
    for o in (dict(relTol=1e-12, maxNumIterations=2, factMethod=0),
              dict(relTol=1e-14, factMethod=0)):
        Xs, its, codes = one_by_one(sap, A, o, P, B, X0)
        Xb, itb, codeb, ok = batched(sap, A, o, P, B, X0)
        np.testing.assert_array_equal(Xb, Xs)
        assert itb == max(its) and codeb == codes[0] and not ok
        assert codeb in (-1, -3)


def test_batched_refinement_reaches_tolerance(sap, oracle):
    n, k, P = 40000, 60, 16
    band = oracle.gen_banded(n, k, 0.8, 331)
    A = sap.BandedMatrix(dev(band))
    nr = 8
    B = np.stack([oracle.gen_vector(n, 340 + c) for c in range(nr)], axis=1)
    X0 = np.zeros_like(B)
    o = dict(relTol=1e-10, refine=8, factMethod=0)
    Xs, its, codes = one_by_one(sap, A, o, P, B, X0)
    Xb, itb, codeb, ok = batched(sap, A, o, P, B, X0)
    assert ok and codeb == 1
    for c in range(nr):
        r = B[:, c] - oracle.band_spmv(band, Xb[:, c])
        assert np.linalg.norm(r) / np.linalg.norm(B[:, c]) <= 1e-10
        np.testing.assert_allclose(Xb[:, c], Xs[:, c], rtol=0, atol=1e-8 * np.abs(Xs[:, c]).max())
    assert abs(itb - max(its)) <= 1.0


def test_batched_graph_replays_match_eager(sap, oracle):
    """Two solves in a row on one solver with the graph (the second replays a
    freshly captured graph), each equal to the eager batched solve."""
    n, k, P = 8000, 16, 4
    band = oracle.gen_banded(n, k, 0.8, 351)
    A = sap.BandedMatrix(dev(band))
    B = np.stack([oracle.gen_vector(n, 352 + c) for c in range(6)], axis=1)
    X0 = np.zeros_like(B)
    eager, _, _, _ = batched(sap, A, dict(relTol=1e-8), P, B, X0, graph=0)
    g = sap.Solver(P, sap.Options(batch=1, graph=1, relTol=1e-8))
    g.setup(A)
    for rep in range(2):
        X = dev(X0)
        g.solve(None, dev(B), X)
        np.testing.assert_array_equal(host(X), eager)


SYSTEM_SETS = [
    # SaP-C systems, one diagonal system: per-system launches
    ([(9000, 20, 4), (5000, 7, 2), (12000, 45, 6), (3000, 0, 1)], 0),
    # Block (SaP-D) systems: one fused sweep launch for all of them
    ([(9000, 20, 4), (5000, 7, 2), (12000, 45, 6), (7000, 130, 3), (4000, 230, 2)], 1),
]


@pytest.mark.parametrize("specs,block", SYSTEM_SETS)
@pytest.mark.parametrize("fuse", [1, 0])
def test_system_batch_equals_single_systems_bitwise(sap, oracle, specs, block, fuse):
    """Several independent banded systems of different sizes and widths
    (SystemBatch, sap_batch_*): every column equals solving it alone with
    its own system's solver, bit for bit (no refinement), with the systems'
    sweeps in one launch (fuse, Block systems) or one launch per system."""
    mats, bands, Bs = [], [], []
    for i, (n, k, P) in enumerate(specs):
        band = gen_band(oracle, n, k, 1.2, 400 + i)
        bands.append(band)
        mats.append(sap.BandedMatrix(dev(band)))
        Bs.append(np.stack([oracle.gen_vector(n, 410 + 10 * i + c) for c in range(3)], axis=1))
    opts = dict(relTol=1e-6, precondType=sap.Block if block else sap.Spike, factMethod=sap.LU_UL)   # FP32, no refinement
    sb = sap.SystemBatch([P for _, _, P in specs], sap.Options(fuse=fuse, **opts))
    sb.setup(mats)
    Xd = [torch.zeros((n, 3), dtype=torch.float64, device="cuda") for n, _, _ in specs]
    ok = sb.solve([dev(B) for B in Bs], Xd)
    res = sb.results()
    assert len(res) == 3 * len(specs) and ok == all(r[0] == 1 for r in res)
    for i, (n, k, P) in enumerate(specs):
        Xs, its, codes = one_by_one(sap, mats[i], opts, P, Bs[i], np.zeros_like(Bs[i]))
        np.testing.assert_array_equal(host(Xd[i]), Xs)
        assert [r[1] for r in res[3 * i:3 * i + 3]] == its


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("block", [0, 1])
def test_autotuned_launches_give_the_same_bits(sap, oracle, prec, block):
    """options.autotune = 1 times the conversion tile widths and the sweep
    staging depths on the first setup of a half-bandwidth and keeps the
    fastest; neither changes the arithmetic, so factors, M^{-1} v and the
    solution equal the untuned run bit for bit, and the choices are listed
    (for the SaP-C and the block-Jacobi preconditioners)."""
    n, k, P = 20000, 100 + prec - 4 * block, 8
    band = oracle.gen_banded(n, k, 0.8 + block, 501)
    A = sap.BandedMatrix(dev(band))
    b = dev(oracle.gen_vector(n, 502))
    out = []
    for tune in (0, 1):
        g = sap.Solver(P, sap.Options(factMethod=sap.LU_UL, relTol=1e-10, refine=6, autotune=tune,
                                      precValueType=prec, precondType=sap.Block if block else sap.Spike))
        g.setup(A)
        z = torch.empty(n, dtype=torch.float64, device="cuda")
        g.precond_apply(b, z)
        x = torch.zeros_like(b)
        assert g.solve(None, b, x)
        out.append((g.factors(), host(z), host(x)))
    for a, c in zip(out[0], out[1]):
        np.testing.assert_array_equal(a, c)
    ent = [e for e in sap.autotune_entries() if e["k"] == k and e["elem_bytes"] == (8 if prec else 4)]
    fams = {e["family"] for e in ent}
    assert fams == {0, 1}
    for e in ent:
        assert e["candidates"] >= 2 and e["best_ms"] > 0 and e["best_ms"] <= e["default_ms"] + 1e-9
