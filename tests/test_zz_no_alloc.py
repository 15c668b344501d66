"""sap_solve_noalloc (the solve whose Krylov vectors live in the library's
static device array) against sap_solve through the C ABI: two calls in a
row on one GPU, the second with a smaller system, each bit-identical to
sap_solve on a fresh solver of the same setup (same kernels in the same
order).  Also the size limit and the profiling restriction."""
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


def system(sap, n, k, dominance, seed):
    A = sap.BandedMatrix.generate(n, k, dominance, seed)
    xt = sap.gen_vector(n, seed + 1)
    b = torch.empty_like(xt)
    sap.band_spmv(A, xt, b)
    return A, b


def run(sap, A, b, P, opts, noalloc):
    s = sap.Solver(P, sap.Options(**opts))
    s.setup(A)
    x = torch.zeros_like(b)
    ok = (s.solve_noalloc if noalloc else s.solve)(None, b, x)
    torch.cuda.synchronize()
    st = s.getStats()
    return ok, s.getMonitorCode(), st.numIterations, st.relResidualNorm, x.cpu().numpy()


def test_noalloc_solve_equals_solve_twice_then_smaller(sap):
    big = dict(precondType=sap.Spike, factMethod=sap.LU_UL, relTol=1e-10, refine=6)
    small = dict(big, refine=0, relTol=1e-6)   # the reference single solve, FP32 factors
    cases = [(system(sap, sap.NOALLOC_MAX_N, 40, 0.8, 5), 16, big),
             (system(sap, 20000, 25, 1.0, 9), 8, small)]
    expect = [run(sap, A, b, P, o, noalloc=False) for (A, b), P, o in cases]
    # two allocation-free calls in a row; the second, smaller, finds the
    # first call's contents in the array
    for ((A, b), P, o), ref in zip(cases, expect):
        print(f"[noalloc] n={A.num_rows} P={P}", flush=True)
        got = run(sap, A, b, P, o, noalloc=True)
        assert got[:4] == ref[:4]
        assert np.array_equal(got[4], ref[4])
        assert got[0] and got[1] == 1 and got[3] <= o["relTol"]


def test_noalloc_several_right_hand_sides(sap):
    n, k = 30000, 12
    A = sap.BandedMatrix.generate(n, k, 1.5, 3)
    B = torch.stack([sap.gen_vector(n, 40 + c) for c in range(3)], dim=1)
    xs = []
    for noalloc in (True, False):
        # batch=0: sap_solve takes the columns one after another, as the
        # allocation-free solve does
        s = sap.Solver(4, sap.Options(precondType=sap.Block, relTol=1e-10, refine=4, batch=0))
        s.setup(A)
        X = torch.zeros_like(B)
        assert (s.solve_noalloc if noalloc else s.solve)(None, B, X)
        xs.append(X.cpu().numpy())
    assert np.array_equal(xs[0], xs[1])


def test_noalloc_limits(sap):
    n = sap.NOALLOC_MAX_N + 1
    A = sap.BandedMatrix.generate(n, 3, 2.0, 1)
    s = sap.Solver(2, sap.Options(precondType=sap.Block))
    s.setup(A)
    b = sap.gen_vector(n, 2)
    x = torch.zeros_like(b)
    with pytest.raises(sap.SapError) as e:
        s.solve_noalloc(None, b, x)
    assert e.value.reason == -8 and "SAP_NOALLOC_MAX_N" in str(e.value)
    small = sap.BandedMatrix.generate(1000, 3, 2.0, 1)
    p = sap.Solver(2, sap.Options(precondType=sap.Block, profile=1))
    p.setup(small)
    b = sap.gen_vector(1000, 2)
    with pytest.raises(sap.SapError) as e:
        p.solve_noalloc(None, b, torch.zeros_like(b))
    assert e.value.reason == -9
    # before setup: Illegal_solve, as sap_solve
    q = sap.Solver(2, sap.Options())
    with pytest.raises(sap.SapError) as e:
        q.solve_noalloc(None, b, torch.zeros_like(b))
    assert e.value.reason == -4
