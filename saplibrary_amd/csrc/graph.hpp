// Host-side reordering pipeline for CSR input (DB, RCM, drop-off); see
// graph.hip.
#pragma once

#include "sap_internal.hpp"

#include <vector>

namespace sap_amd {

struct GraphOptions {
    bool doDB = true;              // performDB
    bool scale = true;             // applyScaling
    bool dbFirstStageOnly = false;
    bool doRCM = true;             // performReorder && maxBandwidth > 64 (sap/precond.h:1726-1728)
    int maxBandwidth = 0x7fffffff;
    double dropOffFrac = 0;
    bool testDB = false;           // testDB: dominance statistics, then stop after DB
};

template <class T>
struct Reordering {
    int n = 0;
    int k_db = 0;                  // half-bandwidth after DB (Stats::bandwidthDB)
    int k_reorder = 0;             // after RCM (Stats::bandwidthReorder)
    int k = 0;                     // after drop-off (Stats::bandwidth)
    double dropoff_actual = 0;     // Stats::actualDropOff
    std::vector<int> db_perm;      // DB row permutation: original row -> row
    std::vector<int> row_perm;     // final row permutation (DB then RCM), m_optPerm
    std::vector<int> col_reord;    // RCM reordering: new index -> original, m_optReordering
    std::vector<T> row_scale, col_scale;   // DB scalings (1 without DB)
    // reordered (and dropped-off) matrix, CSR, values in T
    std::vector<int> rp, ci;
    std::vector<T> val;
    // for value updates (the reference's tracked maps, sap/graph.h:1216-1361):
    // entry_orig[e] = index in the input arrays of reordered entry e;
    // entry_scale[j] = DB scale factor of input entry j (empty: all 1)
    std::vector<int> entry_orig;
    std::vector<T> entry_scale;
    // host wall times, milliseconds
    double time_DB = 0, time_reorder = 0, time_dropOff = 0;
    // testDB (sap/graph.h:468-490, :537-561): d_i = |a_ii| / sum_{j != i}
    // |a_ij| sorted; its minimum and its entry n/100, before and after DB
    double diag_dom_ori = 0, d_p1_ori = 0, diag_dom = 0, d_p1 = 0;
};

// The variable-bandwidth assembly of a reordered matrix into P partitions
// (transformToBandedMatrix's variable-bandwidth branch, sap/precond.h:
// 1785-1800, and Graph::assembleOffDiagMatrices, secondLevelReordering +
// partitionedRCM, assembleBandedMatrix, sap/graph.h:756-1206, :1593-1757).
template <class T>
struct VarBand {
    int n = 0, k = 0, P = 1;
    // coupling blocks: WV holds them with compressed columns (the spike
    // right-hand sides), offDiags as they are; 2 k^2 per boundary
    std::vector<T> WV, offDiags;
    std::vector<int> widths_left, widths_right;   // used columns per boundary
    std::vector<int> perms_left, perms_right;     // (P-1) k column maps, -1 unused
    std::vector<char> exists;                     // rows with a coupling entry to the right
    std::vector<int> second_reorder, second_perm; // per-partition RCM (new -> old, old -> new)
    std::vector<int> first_rows;                  // P-1
    std::vector<int> ks;                          // per-partition half-bandwidths
    std::vector<int64_t> boffsets;                // P+1 band offsets
    std::vector<int> ks_col, ks_row;              // per column / row extents
    // diagonal-block entries after the second-level permutation (CSR in T)
    std::vector<int> drp, dci;
    std::vector<T> dval;
};

// rp/ci/val: the reordered, dropped-off matrix (Reordering::rp/ci/val), k its
// half-bandwidth, P the partition count (already clamped, > 1).
template <class T>
VarBand<T> assemble_var(int n, int k, int P, const std::vector<int>& rp, const std::vector<int>& ci,
                        const std::vector<T>& val);

// Sums over the per-row / per-column band extents ks_row, ks_col of the
// assembled band (Graph::assembleBandedMatrix, sap/graph.h:1019-1091 for one
// partition, :1093-1206 over the diagonal blocks for several), from which
// Solver::setup derives Stats::nuKf and Stats::flops_LU (sap/solver.h:358-399).
struct BandProfile {
    double sum_row = 0, sum_col = 0, sum_prod = 0;
};
BandProfile band_profile(int n, const int* rp, const int* ci, int P);

// Throws Error(SAP_ERR_SINGULAR) when DB finds no perfect matching and
// Error(SAP_ERR_NEG_DB_WEIGHT) on a negative reduced cost.
template <class T>
Reordering<T> reorder_csr(int n, const int* rp, const int* ci, const double* val, const GraphOptions& o);

}  // namespace sap_amd
