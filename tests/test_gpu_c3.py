"""The batching workload's two parts on the GPU at reduced sizes, through the
C ABI (the full sizes are bench.py --workload c3):

  * uneven banded systems (N ~ U[10k, 60k], K ~ U[10, 150], diagonal 1.2x
    the off-diagonal abs-sum, the workload's generator and seeds) with 4
    right-hand sides each, SaP-D P=16, FP32 factors and FP64 refinement, all
    columns in one batched solve: every column reaches a true relative
    residual <= 1e-10 (the library's banded SpMV on the input matrix), and
    equals the one-system, one-column solve to the tolerance;
  * an update() sequence on one banded pattern (N=100k, K=120, 16 steps,
    values x (1 + U[-0.05, 0.05]) per step), each step followed by a solve
    to 1e-10; after the last step the stored factors equal a fresh setup of
    the last values bit for bit (update() refactors the same way).
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


OPTS = dict(factMethod=0, precValueType=0, relTol=1e-10, refine=8)


def test_c3_batched_systems_reach_tolerance(sap):
    rng = np.random.default_rng(30)
    specs = [(int(n), int(k), 30 + i) for i, (n, k) in
             enumerate(zip(rng.integers(10_000, 60_001, 8), rng.integers(10, 151, 8)))]
    print(f"[c3] systems {specs}", flush=True)
    mats = [sap.BandedMatrix.generate(n, k, 1.2, seed) for n, k, seed in specs]
    B = [torch.stack([sap.gen_vector(n, 1000 + 16 * seed + r) for r in range(4)], dim=1) for n, k, seed in specs]
    o = dict(precondType=sap.Block, **OPTS)
    sb = sap.SystemBatch([16] * len(specs), sap.Options(**o))
    sb.setup(mats)
    X = [torch.zeros_like(b) for b in B]
    assert sb.solve(B, X)
    for (n, k, seed), A, b, x in zip(specs, mats, B, X):
        s = sap.Solver(16, sap.Options(batch=0, **o))
        s.setup(A)
        for r in range(4):
            y = torch.empty(n, dtype=torch.float64, device="cuda")
            sap.band_spmv(A, x[:, r].contiguous(), y)
            rel = float(torch.linalg.norm(b[:, r] - y) / torch.linalg.norm(b[:, r]))
            assert rel <= 1e-10, (n, k, r, rel)
            x1 = torch.zeros(n, dtype=torch.float64, device="cuda")
            assert s.solve(None, b[:, r].contiguous(), x1)
            err = float(torch.linalg.norm(x1 - x[:, r]) / torch.linalg.norm(x1))
            assert err <= 1e-8, (n, k, r, err)


def test_c3_update_sequence(sap):
    n, k, P = 100_000, 120, 16
    base = sap.BandedMatrix.generate(n, k, 1.2, 40)
    xt = sap.gen_vector(n, 41)
    o = dict(precondType=sap.Block, trackReordering=1, **OPTS)
    s = sap.Solver(P, sap.Options(**o))
    s.setup(base)
    g = torch.Generator(device="cuda")
    g.manual_seed(42)
    vals = base.band.clone()
    b = torch.empty_like(xt)
    for step in range(16):
        vals = vals * (1.0 + (torch.rand(vals.shape, generator=g, device="cuda", dtype=torch.float64) - 0.5) * 0.1)
        A = sap.BandedMatrix(vals)
        sap.band_spmv(A, xt, b)
        assert s.update(vals)
        x = torch.zeros_like(b)
        assert s.solve(None, b, x), (step, s.getMonitorMessage())
        y = torch.empty_like(b)
        sap.band_spmv(A, x, y)
        rel = float(torch.linalg.norm(b - y) / torch.linalg.norm(b))
        assert rel <= 1e-10, (step, rel)
    fresh = sap.Solver(P, sap.Options(**o))
    fresh.setup(sap.BandedMatrix(vals))
    np.testing.assert_array_equal(s.factors(), fresh.factors())
