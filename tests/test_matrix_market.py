"""Matrix Market input (sap_mm_read), the on-disk format the reference's
drivers read (examples/matrix_market/driver_mm.cu:135-140).

The reader is host code, so these tests run without a GPU.  They are pinned
by scipy's independent Matrix Market implementation (scipy.io.mmwrite writes
the files, scipy.io.mmread's matrix is the expected result) and by files
written by hand for the cases scipy does not produce (comments between
entries, duplicates, malformed input)."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp


def _expected_csr(A):
    """CSR of a scipy matrix, sorted by (row, column), explicit zeros dropped."""
    A = sp.csr_matrix(A)
    A.eliminate_zeros()
    A.sort_indices()
    return A.indptr, A.indices, A.data


def _check(sap, path, A):
    rows, cols, rp, ci, val = sap.read_matrix_market(str(path))
    assert (rows, cols) == A.shape
    erp, eci, eval_ = _expected_csr(A)
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(ci, eci)
    np.testing.assert_array_equal(val, eval_)   # decimal text -> double: exact both ways


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["real", "integer"])
def test_coordinate_files_match_scipy(sap, tmp_path, symmetry, field):
    rng = np.random.default_rng(5)
    n = 40
    A = sp.random(n, n, density=0.15, random_state=7, format="coo").toarray()
    if field == "integer":
        A = np.round(A * 20 - 10)
    if symmetry == "symmetric":
        A = np.tril(A) + np.tril(A, -1).T
    elif symmetry == "skew-symmetric":
        A = np.tril(A, -1) - np.tril(A, -1).T
    A = sp.coo_matrix(A)
    path = tmp_path / "a.mtx"
    scipy.io.mmwrite(str(path), A, field=field, symmetry=symmetry, precision=17)
    assert f"coordinate {field} {symmetry}" in path.read_text().splitlines()[0]
    expected = scipy.io.mmread(str(path))
    _check(sap, path, expected)
    assert rng is not None


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
def test_array_files_match_scipy(sap, tmp_path, symmetry):
    rng = np.random.default_rng(11)
    n = 9
    A = rng.standard_normal((n, n))
    A[rng.random((n, n)) < 0.3] = 0.0
    if symmetry == "symmetric":
        A = np.tril(A) + np.tril(A, -1).T
    elif symmetry == "skew-symmetric":
        A = np.tril(A, -1) - np.tril(A, -1).T
    path = tmp_path / "d.mtx"
    scipy.io.mmwrite(str(path), A, symmetry=symmetry, precision=17)
    assert "array" in path.read_text().splitlines()[0]
    _check(sap, path, np.asarray(scipy.io.mmread(str(path))))


def test_pattern_comments_duplicates_and_rectangular(sap, tmp_path):
    path = tmp_path / "p.mtx"
    path.write_text(
        "%%MatrixMarket matrix coordinate pattern general\n"
        "% a comment\n"
        "\n"
        "3 4 5\n"
        "3 4\n"
        "% a comment between entries\n"
        "1 1\n"
        "2 3\n"
        "1 1\n"
        "3 1\n")
    rows, cols, rp, ci, val = sap.read_matrix_market(str(path))
    assert (rows, cols) == (3, 4)
    # sorted by (row, column); the duplicate (1, 1) is kept as two entries
    np.testing.assert_array_equal(rp, [0, 2, 3, 5])
    np.testing.assert_array_equal(ci, [0, 0, 2, 0, 3])
    np.testing.assert_array_equal(val, np.ones(5))


def test_duplicates_keep_file_order(sap, tmp_path):
    path = tmp_path / "dup.mtx"
    path.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n2 2 5.0\n1 2 -1.5\n2 2 7.0\n")
    _, _, rp, ci, val = sap.read_matrix_market(str(path))
    np.testing.assert_array_equal(rp, [0, 1, 3])
    np.testing.assert_array_equal(ci, [1, 1, 1])
    np.testing.assert_array_equal(val, [-1.5, 5.0, 7.0])


@pytest.mark.parametrize("text,reason", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", "complex"),
    ("%%MatrixMarket matrix coordinate real hermitian\n1 1 1\n1 1 1\n", "hermitian"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", "missing"),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", "out of range"),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 2 1\n1 2 1\n", "lower triangle"),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n", "short"),
    ("not a matrix market file\n", "banner"),
])
def test_malformed_files_raise(sap, tmp_path, text, reason):
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    with pytest.raises(sap.SapError) as e:
        sap.read_matrix_market(str(path))
    assert e.value.reason == -8 and reason in str(e.value)


def test_missing_file_raises(sap, tmp_path):
    with pytest.raises(sap.SapError):
        sap.read_matrix_market(str(tmp_path / "nope.mtx"))


def test_sparse_workload_round_trip(sap, tmp_path):
    """A matrix of the sparse workload's kind written by scipy reads back as
    the CSR the solver's setup takes."""
    from tests.csr_cases import sparse_band
    rp, ci, val = sparse_band(600, 40, 5, seed=3)
    A = sp.csr_matrix((val, ci, rp), shape=(600, 600))
    path = tmp_path / "w.mtx"
    scipy.io.mmwrite(str(path), A, precision=17)
    _check(sap, path, A)
