"""The learned option selector (saplibrary_amd/selector.py).

CPU: the features of a banded matrix equal those of the same matrix given
as CSR (except the format flag); the diagonal-dominance features are the
testDB statistics that the shipped host reordering computes
(sap_reordering_dominance: minimum and 1st percentile); the targets are
log time ratios to the defaults with non-converged option sets penalised.

GPU (the network trained by the library's kernels on the committed corpus
of solver-measured timings, tests/golden/selector_corpus.npz made by
tests/golden/make_selector_corpus.py): on the held-out fifth of the
corpus the chosen options are no slower than the defaults on at least 90%
of the matrices (the fallback to the defaults charged when a choice does
not converge), and faster in the geometric mean.
"""
import math
import os

import numpy as np
import pytest

from saplibrary_amd import selector as S
from saplibrary_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORPUS = os.path.join(ROOT, "tests", "golden", "selector_corpus.npz")


def _band_to_csr(band):
    n, w = band.shape
    k = (w - 1) // 2
    rp, ci, val = [0], [], []
    for i in range(n):
        for j in range(max(0, i - k), min(n, i + k + 1)):
            v = band[i, k + j - i]
            if v != 0:
                ci.append(j)
                val.append(v)
        rp.append(len(ci))
    return np.array(rp), np.array(ci), np.array(val)


def test_band_and_csr_features_agree(oracle):
    band = oracle.gen_banded(3000, 17, 0.9, 5)
    fb = S.features_band(band)
    fc = S.features_csr(*_band_to_csr(band))
    assert fb[0] == 0 and fc[0] == 1
    np.testing.assert_allclose(fb[1:], fc[1:], rtol=1e-12, atol=1e-12)


@pytest.fixture(scope="module")
def sap():
    import saplibrary_amd
    if not os.path.exists(saplibrary_amd.LIB_PATH):
        pytest.skip("libsap_amd.so not built")
    saplibrary_amd.lib()
    return saplibrary_amd


def test_dominance_features_are_the_testdb_statistics(sap):
    rp, ci, val = W.general_sparse(5000, 80, 6, seed=3, weak_seed=4)
    f = S.features_csr(rp, ci, val)
    r = sap.reorder_csr(rp, ci, val, sap.Options(testDB=1, precValueType=sap.PREC_DOUBLE))
    assert f[8] == pytest.approx(math.log10(1e-8 + r["diag_dom_ori"]), abs=1e-12)
    assert f[9] == pytest.approx(math.log10(1e-8 + r["d_p1_ori"]), abs=1e-12)


def test_targets_penalise_failures():
    T = np.array([[10.0] * len(S.GRID)])
    OK = np.ones_like(T, dtype=bool)
    T[0, 3] = 5.0
    OK[0, 7] = False
    Y, M = S.targets(T, OK)
    assert Y[0, S.DEFAULT_INDEX] == 0 and Y[0, 3] == pytest.approx(math.log10(0.5))
    assert Y[0, 7] == pytest.approx(math.log10(100.0 / 10.0)) and M.all()
    Y, M = S.targets(T, np.zeros_like(OK))
    assert not M.any()


torch = pytest.importorskip("torch")


@pytest.mark.gpu
def test_learned_options_beat_defaults_on_held_out(sap):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(CORPUS), "tests/golden/selector_corpus.npz is missing"
    z = np.load(CORPUS)
    F, T, OK = z["features"], z["times"], z["ok"]
    assert int(z["default_index"]) == S.DEFAULT_INDEX and len(z["grid"]) == len(S.GRID)
    keep = OK[:, S.DEFAULT_INDEX] | OK.any(axis=1)
    F, T, OK = F[keep], T[keep], OK[keep]
    m = len(F)
    # five-fold cross-validation: every matrix is judged by an ensemble that
    # did not see it (one fold of ~30 matrices is too few to judge alone)
    ratios = []
    for fold in range(5):
        held = np.arange(m) % 5 == fold
        sel = S.Selector(sap, torch, seed=1)
        losses = sel.fit(F[~held], T[~held], OK[~held], epochs=3000)
        assert losses[-1] < 0.5 * losses[0]
        # only matrices the defaults solve can be compared with them
        ho = held & OK[:, S.DEFAULT_INDEX]
        r = S.evaluate(sel.choose(F[ho]), T[ho], OK[ho])
        print(f"[selector] fold {fold}: {ho.sum()} held-out matrices, no slower on {np.mean(r <= 1.0):.2%}",
              flush=True)
        ratios.append(r)
    ratio = np.concatenate(ratios)
    print(f"[selector] cross-validated over {len(ratio)} matrices: no slower on {np.mean(ratio <= 1.0):.2%}, "
          f"geometric-mean time ratio {np.exp(np.mean(np.log(ratio))):.3f}")
    assert np.mean(ratio <= 1.0) >= 0.9
    assert np.exp(np.mean(np.log(ratio))) < 1.0
