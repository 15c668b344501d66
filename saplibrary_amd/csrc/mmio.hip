// Matrix Market reader (host only): the on-disk input of the reference's
// drivers, which load a CSR matrix with cusp::io::read_matrix_market_file
// (examples/matrix_market/driver_mm.cu:135-140) before Solver::setup.
//
// CUSP is not part of the reference tree; the behaviour restated here is the
// Matrix Market format's own (NIST MM exchange format, 1996) as CUSP's
// reader applies it:
//   * banner "%%MatrixMarket matrix <coordinate|array> <real|integer|pattern>
//     <general|symmetric|skew-symmetric>" (case-insensitive); complex and
//     hermitian matrices are rejected (the solver is real);
//   * '%' lines are comments; the size line is "rows cols nnz" (coordinate)
//     or "rows cols" (array);
//   * coordinate entries are 1-based "i j [value]", pattern entries get 1.0;
//     symmetric files hold one triangle: every off-diagonal (i, j) also
//     yields (j, i), negated for skew-symmetric;
//   * array files are dense column-major (symmetric / skew: the lower
//     triangle, column by column, skew without the zero diagonal); zeros are
//     not kept as entries;
//   * the result is sorted by (row, column); duplicate coordinates stay
//     separate entries in file order (no summation), as a COO -> CSR
//     conversion leaves them.
#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sap_c.h"

struct sap_mm_matrix {
    int rows = 0, cols = 0;
    std::vector<int> rp, ci;
    std::vector<double> val;
    std::string error;
};

namespace {

std::string lower(std::string s)
{
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

// false on a malformed file; m.error says why
bool parse(std::istream& in, sap_mm_matrix& m)
{
    std::string line;
    if (!std::getline(in, line)) { m.error = "empty file"; return false; }
    std::istringstream banner(line);
    std::string tag, object, format, field, symmetry;
    banner >> tag >> object >> format >> field >> symmetry;
    if (tag != "%%MatrixMarket") { m.error = "missing %%MatrixMarket banner"; return false; }
    object = lower(object); format = lower(format); field = lower(field); symmetry = lower(symmetry);
    if (object != "matrix") { m.error = "object is not 'matrix'"; return false; }
    const bool coordinate = format == "coordinate";
    if (!coordinate && format != "array") { m.error = "format must be coordinate or array"; return false; }
    if (field == "complex") { m.error = "complex matrices are not supported"; return false; }
    const bool pattern = field == "pattern";
    if (!pattern && field != "real" && field != "integer" && field != "double") {
        m.error = "unknown field '" + field + "'";
        return false;
    }
    if (pattern && !coordinate) { m.error = "pattern needs the coordinate format"; return false; }
    if (symmetry == "hermitian") { m.error = "hermitian matrices are not supported"; return false; }
    const bool symmetric = symmetry == "symmetric";
    const bool skew = symmetry == "skew-symmetric";
    if (!symmetric && !skew && symmetry != "general") {
        m.error = "unknown symmetry '" + symmetry + "'";
        return false;
    }
    // size line (after comments and blank lines)
    while (std::getline(in, line)) {
        size_t p = line.find_first_not_of(" \t\r");
        if (p == std::string::npos || line[p] == '%') continue;
        break;
    }
    long long rows = -1, cols = -1, entries = -1;
    {
        std::istringstream sz(line);
        sz >> rows >> cols;
        if (coordinate) sz >> entries;
        if (!sz || rows < 0 || cols < 0 || (coordinate && entries < 0) || rows > INT32_MAX || cols > INT32_MAX) {
            m.error = "bad size line";
            return false;
        }
    }
    if ((symmetric || skew) && rows != cols) { m.error = "a symmetric matrix must be square"; return false; }
    std::vector<long long> I, J;
    std::vector<double> V;
    auto push = [&](long long i, long long j, double v) {
        I.push_back(i); J.push_back(j); V.push_back(v);
        if (i != j && (symmetric || skew)) {
            I.push_back(j); J.push_back(i); V.push_back(skew ? -v : v);
        }
    };
    // every token after the size line, in order
    auto next_token = [&](std::string& tok) -> bool {
        while (true) {
            int c = in.peek();
            if (c == EOF) return false;
            if (std::isspace(c)) { in.get(); continue; }
            if (c == '%') { std::getline(in, tok); continue; }
            in >> tok;
            return (bool)in;
        }
    };
    std::string t;
    auto read_num = [&](double& v) -> bool {
        if (!next_token(t)) return false;
        char* end = nullptr;
        v = std::strtod(t.c_str(), &end);
        return end && *end == '\0';
    };
    if (coordinate) {
        I.reserve(entries); J.reserve(entries); V.reserve(entries);
        for (long long e = 0; e < entries; ++e) {
            double di, dj, v = 1.0;
            if (!read_num(di) || !read_num(dj) || (!pattern && !read_num(v))) {
                m.error = "entry " + std::to_string(e + 1) + " is missing or malformed";
                return false;
            }
            const long long i = (long long)di, j = (long long)dj;
            if ((double)i != di || (double)j != dj || i < 1 || i > rows || j < 1 || j > cols) {
                m.error = "entry " + std::to_string(e + 1) + " has an index out of range";
                return false;
            }
            if ((symmetric || skew) && j > i) {
                m.error = "a symmetric file must hold the lower triangle";
                return false;
            }
            if (skew && i == j) { m.error = "a skew-symmetric file has no diagonal"; return false; }
            push(i - 1, j - 1, v);
        }
    } else {
        for (long long j = 0; j < cols; ++j) {
            const long long i0 = symmetric ? j : skew ? j + 1 : 0;
            for (long long i = i0; i < rows; ++i) {
                double v;
                if (!read_num(v)) { m.error = "array data is short"; return false; }
                if (v != 0.0) push(i, j, v);
            }
        }
    }
    // sort by (row, column), keeping file order among equal coordinates
    std::vector<size_t> ord(I.size());
    std::iota(ord.begin(), ord.end(), (size_t)0);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) {
        return I[a] != I[b] ? I[a] < I[b] : J[a] < J[b];
    });
    if (ord.size() > (size_t)INT32_MAX) { m.error = "too many entries for int32 CSR"; return false; }
    m.rows = (int)rows;
    m.cols = (int)cols;
    m.rp.assign((size_t)rows + 1, 0);
    m.ci.resize(ord.size());
    m.val.resize(ord.size());
    for (size_t e = 0; e < ord.size(); ++e) {
        m.rp[(size_t)I[ord[e]] + 1]++;
        m.ci[e] = (int)J[ord[e]];
        m.val[e] = V[ord[e]];
    }
    for (long long i = 0; i < rows; ++i) m.rp[(size_t)i + 1] += m.rp[(size_t)i];
    return true;
}

}  // namespace

extern "C" {

sap_mm_matrix* sap_mm_read(const char* path, int* err, char* msg, int msg_len)
{
    auto fail = [&](int c, sap_mm_matrix* m, const std::string& why) {
        if (err) *err = c;
        if (msg && msg_len > 0) {
            std::strncpy(msg, why.c_str(), (size_t)msg_len - 1);
            msg[msg_len - 1] = '\0';
        }
        delete m;
        return (sap_mm_matrix*)nullptr;
    };
    if (!path) return fail(SAP_ERR_BADARG, nullptr, "null path");
    auto* m = new (std::nothrow) sap_mm_matrix();
    if (!m) return fail(SAP_ERR_OOM, nullptr, "out of host memory");
    try {
        std::ifstream in(path);
        if (!in) return fail(SAP_ERR_BADARG, m, std::string("cannot open ") + path);
        if (!parse(in, *m)) return fail(SAP_ERR_BADARG, m, m->error);
    } catch (const std::bad_alloc&) {
        return fail(SAP_ERR_OOM, m, "out of host memory");
    }
    if (err) *err = SAP_OK;
    if (msg && msg_len > 0) msg[0] = '\0';
    return m;
}

int sap_mm_info(const sap_mm_matrix* m, int* rows, int* cols, int64_t* nnz)
{
    if (!m) return SAP_ERR_BADARG;
    if (rows) *rows = m->rows;
    if (cols) *cols = m->cols;
    if (nnz) *nnz = (int64_t)m->ci.size();
    return SAP_OK;
}

int sap_mm_get_csr(const sap_mm_matrix* m, int* row_offsets, int* col_indices, double* values)
{
    if (!m) return SAP_ERR_BADARG;
    if (row_offsets) std::copy(m->rp.begin(), m->rp.end(), row_offsets);
    if (col_indices) std::copy(m->ci.begin(), m->ci.end(), col_indices);
    if (values) std::copy(m->val.begin(), m->val.end(), values);
    return SAP_OK;
}

void sap_mm_free(sap_mm_matrix* m) { delete m; }

}  // extern "C"
