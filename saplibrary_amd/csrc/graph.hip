// Host-side reordering pipeline for sparse (CSR) input: diagonal boosting
// (DB, a maximum-product matching with row/column scaling), reverse
// Cuthill-McKee and element drop-off (sap::Graph, sap/graph.h).
//
// This is integer (and, for the DB costs, FP64) host work whose results are
// permutations, scalings and bandwidths; they follow the reference step by
// step so that they are identical to the reference's on the same input,
// including its tie-breaking: the same traversal orders, stable counting
// sorts, and std::priority_queue with the same comparison (libstdc++'s heap
// decides among equal keys exactly as it does for the reference).  The
// matrix values carry the preconditioner type T, as PrecMatrixCsrH does.
#include "graph.hpp"

#include <chrono>

#include <algorithm>
#include <cmath>
#include <queue>
#include <stdexcept>

namespace sap_amd {

namespace {

constexpr double LOC_INFINITY = 1e37;   // Graph<T>::LOC_INFINITY (sap/graph.h:417)

// (node, key) with the reference's CompareValue: the queue's top is the
// smallest key (sap/graph.h:394-398).
template <class K>
struct NodeKey {
    int node;
    K key;
};
template <class K>
struct KeyGreater {
    bool operator()(const NodeKey<K>& a, const NodeKey<K>& b) const { return a.key > b.key; }
};

// Stable counting sort of entries by a new row index: entries of one row
// keep their input order (the reference fills each row from its end while
// walking the entries backwards, sap/graph.h:1334-1347).
// `id` (optional) is carried along with the entries.
template <class T>
void regroup_rows(int n, const std::vector<int>& new_row, std::vector<int>& ci, std::vector<T>& val,
                  std::vector<int>& rp, std::vector<int>* id = nullptr)
{
    const size_t nnz = new_row.size();
    rp.assign(n + 1, 0);
    for (size_t i = 0; i < nnz; ++i) rp[new_row[i] + 1]++;
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    std::vector<int> pos(rp.begin(), rp.end() - 1);
    std::vector<int> ci2(nnz), id2(id ? nnz : 0);
    std::vector<T> v2(nnz);
    for (size_t i = 0; i < nnz; ++i) {
        const int d = pos[new_row[i]]++;
        ci2[d] = ci[i];
        v2[d] = val[i];
        if (id) id2[d] = (*id)[i];
    }
    ci.swap(ci2);
    val.swap(v2);
    if (id) id->swap(id2);
}

// ------------------------------------------------------------------- DB
// Costs c = log(rowmax / |a|) per CSR entry (get_csc_matrix +
// getResidualValues, sap/graph.h:1829-1857, sap/device/db.cuh:12-37); the
// row maxima come from a reduce-by-key over the entries, one per non-empty
// row, in row order.
template <class T>
struct DbState {
    int n;
    const std::vector<int>& rp;
    const std::vector<int>& ci;
    std::vector<double> c;          // reduced-cost base values
    std::vector<double> rowmax;     // per non-empty row, in order
    std::vector<double> u, v;       // duals: u by column index, v by CSR row
    std::vector<int> match;         // match[column] = CSR row
    std::vector<int> rev_entry;     // rev_entry[row] = entry index of its matched edge
    std::vector<char> matched, rev_matched;
    // Dijkstra scratch kept between augmentations (m_DB_*; the per-column
    // part lives in DbCol records during the augmentations)
    std::vector<int> Blist;
    std::vector<int> prev;

    DbState(int n_, const std::vector<int>& rp_, const std::vector<int>& ci_) : n(n_), rp(rp_), ci(ci_) {}
};

template <class T>
void db_costs(DbState<T>& S, const std::vector<T>& val)
{
    const int n = S.n;
    const size_t nnz = val.size();
    S.c.resize(nnz);
    for (size_t j = 0; j < nnz; ++j) S.c[j] = std::fabs((double)val[j]);
    S.rowmax.assign(n, 0.0);
    int run = 0;
    for (int i = 0; i < n; ++i) {
        if (S.rp[i] == S.rp[i + 1]) continue;
        double m = S.c[S.rp[i]];
        for (int j = S.rp[i] + 1; j < S.rp[i + 1]; ++j) m = std::max(m, S.c[j]);
        S.rowmax[run++] = m;
    }
    for (int i = 0; i < n; ++i)
        for (int j = S.rp[i]; j < S.rp[i + 1]; ++j) S.c[j] = std::log(S.rowmax[i] / S.c[j]);
}

// Initial duals and greedy matching (init_reduced_cval, sap/graph.h:1875-2035)
template <class T>
void db_init(DbState<T>& S, bool first_stage_only)
{
    const int n = S.n;
    const auto& rp = S.rp;
    const auto& ci = S.ci;
    auto& c = S.c;
    S.u.assign(n, 0.0);
    {
        int run = 0;   // reduce-by-key minimum, one value per non-empty row
        for (int i = 0; i < n; ++i) {
            if (rp[i] == rp[i + 1]) continue;
            double m = c[rp[i]];
            for (int j = rp[i] + 1; j < rp[i + 1]; ++j) m = std::min(m, c[j]);
            S.u[run++] = m;
        }
    }
    S.v.assign(n, LOC_INFINITY);
    S.match.assign(n, 0);
    S.rev_entry.assign(n, 0);
    S.matched.assign(n, 0);
    S.rev_matched.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        int min_idx = -1;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            if (c[j] > LOC_INFINITY / 2.0) continue;
            const double t = c[j] - S.u[ci[j]];
            if (S.v[i] > t) {
                S.v[i] = t;
                min_idx = j;
            }
        }
        if (min_idx >= 0) {
            const int col = ci[min_idx];
            if (!S.matched[col]) {
                S.rev_matched[i] = 1;
                S.matched[col] = 1;
                S.match[col] = i;
                S.rev_entry[i] = min_idx;
            }
        }
    }
    if (first_stage_only) {
        const double CMP_THRESHOLD = 1e-10;
        for (int i = 0; i < n; ++i) {
            if (S.rev_matched[i]) continue;
            const int cur_v = (int)S.v[i];   // the reference truncates the dual to int here
            bool found = false;
            for (int j = rp[i]; j < rp[i + 1] && !found; ++j) {
                if (c[j] > LOC_INFINITY / 2.0) continue;
                const int col = ci[j];
                if (!S.matched[col]) continue;
                if (c[j] - S.u[col] - cur_v > CMP_THRESHOLD) continue;
                const int i2 = S.match[col];
                for (int j2 = rp[i2]; j2 < rp[i2 + 1]; ++j2) {
                    if (c[j2] > LOC_INFINITY / 2.0) continue;
                    const int col2 = ci[j2];
                    if (S.matched[col2]) continue;
                    S.matched[col2] = 1;
                    S.match[col2] = i2;
                    S.rev_entry[i2] = j2;
                    S.match[col] = i;
                    S.rev_matched[i] = 1;
                    S.rev_entry[i] = j;
                    found = true;
                    break;
                }
            }
        }
        for (int i = 0; i < n; ++i) {
            if (S.rev_matched[i]) continue;
            for (int j = rp[i]; j < rp[i + 1]; ++j) {
                const int col = ci[j];
                if (!S.matched[col]) {
                    S.matched[col] = 1;
                    S.rev_matched[i] = 1;
                    S.match[col] = i;
                    S.rev_entry[i] = j;
                    break;
                }
            }
        }
        for (int i = 0; i < n; ++i) {
            if (S.rev_matched[i]) continue;
            bool found = false;
            for (int j = rp[i]; j < rp[i + 1] && !found; ++j) {
                const int col = ci[j];
                const int i2 = S.match[col];
                for (int j2 = rp[i2]; j2 < rp[i2 + 1]; ++j2) {
                    const int col2 = ci[j2];
                    if (S.matched[col2]) continue;
                    found = true;
                    S.matched[col2] = 1;
                    S.match[col2] = i2;
                    S.rev_entry[i2] = j2;
                    S.match[col] = i;
                    S.rev_matched[i] = 1;
                    S.rev_entry[i] = j;
                    break;
                }
            }
        }
        int last_j = 0;
        for (int i = 0; i < n; ++i) {
            if (S.rev_matched[i]) continue;
            for (int j = last_j; j < n; ++j) {
                if (S.matched[j]) continue;
                S.rev_matched[i] = S.matched[j] = 1;
                S.match[j] = i;
                last_j = j + 1;
                break;
            }
        }
        std::fill(S.u.begin(), S.u.end(), 1.0);
        std::fill(S.v.begin(), S.v.end(), 1.0);
    } else {
        for (int i = 0; i < n; ++i) {
            if (!S.matched[i]) S.u[i] = 0.0;
            if (!S.rev_matched[i]) S.v[i] = 0.0;
        }
    }
}

// Per-column state of the augmenting-path search, one record per column:
// relaxing an edge reads the column's dual, distance and flags together
// (one cache line instead of one per array).  Filled from DbState before
// the augmentations and copied back after them.
struct DbCol {
    double u;        // column dual
    double dval;     // tentative distance (LOC_INFINITY when not reached)
    int match;       // matched CSR row
    int irn;         // entry through which the column was last reached
    char inB, visited, matched, pad;
};

// Shortest augmenting path from CSR row `init` (find_shortest_aug_path,
// sap/graph.h:2045-2173): Dijkstra over reduced costs, then augmentation and
// dual update.  The queue is a binary heap in `heap` driven by
// std::push_heap / std::pop_heap with the reference's comparator, which is
// std::priority_queue's own implementation, so entries leave in the same
// order; the vector is reused across augmentations.
template <class T>
void db_augment(DbState<T>& S, std::vector<DbCol>& C, std::vector<NodeKey<double>>& heap, int init)
{
    const auto& rp = S.rp;
    const auto& ci = S.ci;
    const auto& c = S.c;
    const KeyGreater<double> greater;
    heap.clear();
    double lsp = 0.0, lsap = LOC_INFINITY;
    int cur = init, isap = -1, ksap = -1, b_cnt = 0;
    S.prev[init] = -1;
    while (true) {
        const double vcur = S.v[cur];
        for (int i = rp[cur]; i < rp[cur + 1]; ++i) {
            const int col = ci[i];
            DbCol& cc = C[col];
            if (cc.inB) continue;
            if (c[i] > LOC_INFINITY / 2.0) continue;
            const double reduced = c[i] - cc.u - vcur;
            if (reduced + 1e-10 < 0) throw Error(SAP_ERR_NEG_DB_WEIGHT, "Negative reduced weight in DB.");
            const double d_new = lsp + reduced;
            if (d_new < lsap) {
                if (!cc.matched) {
                    lsap = d_new;
                    isap = col;
                    ksap = i;
                    cc.match = cur;
                } else if (d_new < cc.dval) {
                    cc.dval = d_new;
                    S.prev[cc.match] = cur;
                    heap.push_back({col, d_new});
                    std::push_heap(heap.begin(), heap.end(), greater);
                    cc.irn = i;
                }
            }
        }
        NodeKey<double> top{0, 0.0};
        bool found = false;
        while (!heap.empty()) {
            top = heap.front();
            std::pop_heap(heap.begin(), heap.end(), greater);
            heap.pop_back();
            if (C[top.node].visited) continue;
            found = true;
            break;
        }
        if (!found) break;
        const int t = top.node;
        C[t].visited = 1;
        lsp = top.key;
        if (lsap <= lsp) {
            C[t].visited = 0;
            C[t].dval = LOC_INFINITY;
            break;
        }
        C[t].inB = 1;
        S.Blist[b_cnt++] = t;
        cur = C[t].match;
    }
    if (lsap < LOC_INFINITY / 2.0) {
        C[isap].matched = 1;
        cur = C[isap].match;
        S.v[cur] = c[ksap];
        while (S.prev[cur] >= 0) {
            C[isap].match = cur;
            int next_k = S.rev_entry[cur];
            const int next_i = ci[next_k];
            next_k = C[next_i].irn;
            S.rev_entry[cur] = ksap;
            cur = S.prev[cur];
            isap = next_i;
            ksap = next_k;
        }
        C[isap].match = cur;
        S.rev_entry[cur] = ksap;
        S.rev_matched[cur] = 1;
        for (int i = 0; i < b_cnt; ++i) {
            const int col = S.Blist[i];
            DbCol& cb = C[col];
            const int jv = cb.match;
            const int tk = S.rev_entry[jv];
            cb.u += cb.dval - lsap;
            S.v[jv] = c[tk] - cb.u;
            cb.dval = LOC_INFINITY;
            cb.visited = 0;
            cb.inB = 0;
        }
        // the columns still queued get their distance reset (the order of
        // the resets does not matter)
        for (const auto& q : heap) C[q.node].dval = LOC_INFINITY;
        heap.clear();
    }
}

// Every augmentation of the DB matching (the loop over unmatched rows of
// Graph::DB, sap/graph.h:1256-1263), run on a relabelled copy of the
// matrix for locality: node i (row i and column i of the square pattern)
// gets its position in a breadth-first order of the pattern (restarted at
// the lowest unnumbered node), and the rows are stored in that order with
// each row's entries in their original order.  The search never compares
// node labels (the heap orders by distance, and pushes happen in entry
// order), the rows are still visited in their original order, and the
// same operations run on the same values, so the matching, the duals and
// every tie are those of the original labelling; a randomly permuted
// matrix stops sending every relaxed edge to a random cache line.
template <class T>
void db_augment_all(DbState<T>& S)
{
    const int n = S.n;
    const auto& rp = S.rp;
    const auto& ci = S.ci;
    const size_t nnz = S.c.size();
    // no entries: no row can be matched (the singular check follows)
    if (nnz == 0) return;
    std::vector<int> pos(n, -1), inv(n);
    {
        int head = 0, tail = 0, seed = 0;
        while (tail < n) {
            if (head == tail) {
                while (pos[seed] >= 0) ++seed;
                pos[seed] = tail;
                inv[tail++] = seed;
            }
            const int r = inv[head++];
            for (int e = rp[r]; e < rp[r + 1]; ++e) {
                const int j = ci[e];
                if (pos[j] < 0) {
                    pos[j] = tail;
                    inv[tail++] = j;
                }
            }
        }
    }
    std::vector<int> rp2(n + 1), ci2(nnz), emap(nnz), einv(nnz);
    rp2[0] = 0;
    for (int r2 = 0; r2 < n; ++r2) {
        const int r = inv[r2];
        const int len = rp[r + 1] - rp[r];
        rp2[r2 + 1] = rp2[r2] + len;
        for (int t = 0; t < len; ++t) {
            const int e = rp[r] + t, e2 = rp2[r2] + t;
            ci2[e2] = pos[ci[e]];
            emap[e] = e2;
            einv[e2] = e;
        }
    }
    DbState<T> R(n, rp2, ci2);
    R.c.resize(nnz);
    for (size_t e = 0; e < nnz; ++e) R.c[emap[e]] = S.c[e];
    R.v.resize(n);
    R.rev_entry.resize(n);
    R.rev_matched.resize(n);
    for (int r = 0; r < n; ++r) {
        R.v[pos[r]] = S.v[r];
        R.rev_entry[pos[r]] = emap[S.rev_entry[r]];
        R.rev_matched[pos[r]] = S.rev_matched[r];
    }
    R.Blist.assign(n, 0);
    R.prev.assign(n, 0);
    std::vector<DbCol> cols(n);
    for (int j = 0; j < n; ++j) cols[pos[j]] = {S.u[j], LOC_INFINITY, pos[S.match[j]], 0, 0, 0, S.matched[j], 0};
    std::vector<NodeKey<double>> heap;
    for (int i = 0; i < n; ++i)
        if (!R.rev_matched[pos[i]]) db_augment(R, cols, heap, pos[i]);
    for (int j = 0; j < n; ++j) {
        const DbCol& cj = cols[pos[j]];
        S.u[j] = cj.u;
        S.match[j] = inv[cj.match];
        S.matched[j] = cj.matched;
    }
    for (int r = 0; r < n; ++r) {
        S.v[r] = R.v[pos[r]];
        S.rev_entry[r] = einv[R.rev_entry[pos[r]]];
        S.rev_matched[r] = R.rev_matched[pos[r]];
    }
}

// ------------------------------------------------------------------- RCM
// Symmetric adjacency without self loops, each (i, j) edge inserted as
// (i, j) and (j, i) in entry order, then grouped by row stably
// (buildTopology, sap/graph.h:1766-1809).
void build_topology(int n, const std::vector<int>& rows, const std::vector<int>& cols, std::vector<int>& arp,
                    std::vector<int>& aci)
{
    std::vector<int> r2, c2;
    r2.reserve(rows.size() * 2);
    c2.reserve(rows.size() * 2);
    for (size_t e = 0; e < rows.size(); ++e) {
        const int from = rows[e], to = cols[e];
        if (from == to) continue;
        r2.push_back(from); c2.push_back(to);
        r2.push_back(to); c2.push_back(from);
    }
    std::vector<float> dummy(r2.size());
    regroup_rows(n, r2, c2, dummy, arp);
    aci.swap(c2);
}

int max_edge_length(const std::vector<int>& rows, const std::vector<int>& cols, const int* perm)
{
    int bw = 0;
    for (size_t e = 0; e < rows.size(); ++e) {
        const int a = perm ? perm[rows[e]] : rows[e], b = perm ? perm[cols[e]] : cols[e];
        bw = std::max(bw, std::abs(a - b));
    }
    return bw;
}

// RCM (Graph::RCM, sap/graph.h:1373-1583).  As in the reference, the trial
// loop's level bookkeeping never raises max_level above 0, so at most two
// trials run; the second start vertex is chosen from the vertices at level 0
// by the degrees of vertices 0..count-1.  Returns the bandwidth and fills
// reord (new -> old) and perm (old -> new).
int rcm(int n, const std::vector<int>& rp, const std::vector<int>& ci, std::vector<int>& reord, std::vector<int>& perm)
{
    reord.resize(n);
    perm.resize(n);
    for (int i = 0; i < n; ++i) reord[i] = i;
    std::vector<int> rows(ci.size());
    for (int i = 0; i < n; ++i)
        for (int j = rp[i]; j < rp[i + 1]; ++j) rows[j] = i;
    std::vector<int> deg(n);
    for (int i = 0; i < n; ++i) deg[i] = rp[i + 1] - rp[i];
    int bandwidth = max_edge_length(rows, ci, nullptr);
    std::vector<int> arp, aci;
    build_topology(n, rows, ci, arp, aci);

    const int MAX_NUM_TRIAL = 5, BANDWIDTH_THRESHOLD = 256;
    std::vector<char> tried(n, 0);
    std::vector<int> pushed(n, -1), levels(n, 0), tmp(n), tperm(n);
    const int max_level = 0;
    int p_max_level = 0;
    for (int trial = 0; trial < MAX_NUM_TRIAL; ++trial) {
        std::queue<int> q;
        std::priority_queue<NodeKey<int>, std::vector<NodeKey<int>>, KeyGreater<int>> pq;
        int node;
        if (trial > 0) {
            const int max_pos = (int)(std::max_element(levels.begin(), levels.end()) - levels.begin());
            const int max_count = (int)std::count(levels.begin(), levels.end(), max_level);
            if (max_count > 1) {
                std::vector<int> verts;
                for (int i = 0; i < n; ++i)
                    if (levels[i] == max_level) verts.push_back(i);
                const int pos = (int)(std::min_element(deg.begin(), deg.begin() + max_count) - deg.begin());
                node = verts[pos];
            } else {
                node = max_pos;
            }
            while (tried[node]) node = (node + 1) % n;
        } else {
            node = (int)(std::min_element(deg.begin(), deg.end()) - deg.begin());
        }
        tried[node] = 1;
        levels[node] = 0;
        pushed[node] = trial;
        q.push(node);
        int left = n, j = 0, last = 0;
        while (left--) {
            if (q.empty()) {
                left++;
                int i;
                for (i = last; i < n; ++i) {
                    if (pushed[i] != trial) {
                        q.push(i);
                        pushed[i] = trial;
                        last = i;
                        break;
                    }
                }
                if (i < n) continue;
                return -1;
            }
            node = q.front();
            q.pop();
            tmp[j++] = node;
            const int lev = levels[node];
            for (int e = arp[node]; e < arp[node + 1]; ++e) {
                const int t = aci[e];
                if (pushed[t] != trial) {
                    pushed[t] = trial;
                    pq.push({t, deg[t]});
                    levels[t] = lev + 1;
                }
            }
            while (!pq.empty()) {
                q.push(pq.top().node);
                pq.pop();
            }
        }
        for (int i = 0; i < n; ++i) tperm[tmp[i]] = i;
        const int bw = max_edge_length(rows, ci, tperm.data());
        if (bandwidth > bw) {
            bandwidth = bw;
            reord = tmp;
        }
        if (trial > 0) {
            if (p_max_level >= max_level) break;
            const double ratio = 1.0 * (max_level - p_max_level) / p_max_level;
            if (ratio < 0.01) break;
            p_max_level = max_level;
        } else {
            p_max_level = max_level;
        }
        if (bandwidth <= BANDWIDTH_THRESHOLD) break;
    }
    for (int i = 0; i < n; ++i) perm[reord[i]] = i;
    return bandwidth;
}

// The diagonal-dominance statistics of the testDB mode (Graph::reorder,
// sap/graph.h:468-490 before DB and :537-561 after): per row
// d_i = |a_ii| / sum_{j != i} |a_ij| (in double, from the values in T),
// sorted; the minimum and the entry at n/100.
template <class T>
void dominance_stats(int n, const std::vector<int>& rp, const std::vector<int>& ci, const std::vector<T>& val,
                     double* dmin, double* dp1)
{
    std::vector<double> ds(n, 0.0);
    for (int i = 0; i < n; ++i) {
        double diag = 0.0, sums = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
            if (ci[j] == i) diag = std::fabs(val[j]);
            else sums += std::fabs(val[j]);
        }
        ds[i] = diag / sums;
    }
    std::sort(ds.begin(), ds.end());
    *dmin = ds[0];
    *dp1 = ds[n / 100];
}

}  // namespace

template <class T>
Reordering<T> reorder_csr(int n, const int* rp_in, const int* ci_in, const double* val_in, const GraphOptions& o)
{
    Reordering<T> R;
    R.n = n;
    // (1) values in the preconditioner type, explicit zeros removed
    //     (transformToBandedMatrix, sap/precond.h:1690-1710)
    std::vector<int> rp(n + 1, 0), ci, id;   // id: entry index in the input arrays
    std::vector<T> val;
    ci.reserve(rp_in[n]);
    val.reserve(rp_in[n]);
    id.reserve(rp_in[n]);
    for (int i = 0; i < n; ++i) {
        for (int j = rp_in[i]; j < rp_in[i + 1]; ++j) {
            const T v = (T)val_in[j];
            if (v != (T)0) {
                ci.push_back(ci_in[j]);
                val.push_back(v);
                id.push_back(j);
            }
        }
        rp[i + 1] = (int)ci.size();
    }
    const size_t nnz = ci.size();

    using clock = std::chrono::steady_clock;
    auto since = [](clock::time_point a) {
        return std::chrono::duration<double, std::milli>(clock::now() - a).count();
    };
    const auto t_reorder = clock::now();
    if (o.testDB) dominance_stats(n, rp, ci, val, &R.diag_dom_ori, &R.d_p1_ori);

    // (2) DB (Graph::reorder / Graph::DB, sap/graph.h:448-579, :1216-1361)
    const auto t_db = clock::now();
    R.db_perm.resize(n);
    R.row_scale.assign(n, (T)1);
    R.col_scale.assign(n, (T)1);
    if (o.doDB) {
        DbState<T> S(n, rp, ci);
        db_costs(S, val);
        db_init(S, o.dbFirstStageOnly);
        db_augment_all(S);
        for (int i = 0; i < n; ++i)
            if (!S.matched[i]) throw Error(SAP_ERR_SINGULAR, "Singular matrix found");
        std::vector<double> cs(n), rs(n);
        for (int i = 0; i < n; ++i) {
            cs[i] = std::exp(S.u[i]);
            rs[i] = std::exp(S.v[i]) / S.rowmax[i];
        }
        for (int i = 0; i < n; ++i) R.db_perm[S.match[i]] = i;
        // the matrix is scaled unless scaling is off or only the first stage
        // ran; the scalings are returned either way (the preconditioner
        // applies them to vectors whenever applyScaling is set, as the
        // reference does)
        const bool scale = o.scale && !o.dbFirstStageOnly;
        std::vector<int> new_row(nnz);
        for (int i = 0; i < n; ++i) {
            for (int l = rp[i]; l < rp[i + 1]; ++l) {
                new_row[l] = R.db_perm[i];
                if (scale) val[l] = (T)(rs[i] * cs[ci[l]]) * val[l];
            }
        }
        for (int i = 0; i < n; ++i) {
            R.row_scale[i] = (T)rs[i];
            R.col_scale[i] = (T)cs[i];
        }
        if (scale) {
            // scale factor of every input entry, for value updates
            R.entry_scale.assign(rp_in[n], (T)1);
            for (int i = 0; i < n; ++i)
                for (int l = rp[i]; l < rp[i + 1]; ++l) R.entry_scale[id[l]] = (T)(rs[i] * cs[ci[l]]);
        }
        regroup_rows(n, new_row, ci, val, rp, &id);
        R.time_DB = since(t_db);
    } else {
        for (int i = 0; i < n; ++i) R.db_perm[i] = i;
    }
    {
        int kdb = 0;
        for (int i = 0; i < n; ++i)
            for (int j = rp[i]; j < rp[i + 1]; ++j) kdb = std::max(kdb, std::abs(i - ci[j]));
        R.k_db = kdb;
    }
    if (o.testDB) {
        // the DB test stops here and reports the bandwidth after DB
        // (sap/graph.h:537-561)
        dominance_stats(n, rp, ci, val, &R.diag_dom, &R.d_p1);
        R.k_reorder = R.k = R.k_db;
        R.time_reorder = since(t_reorder);
        R.row_perm = R.db_perm;
        R.col_reord.resize(n);
        for (int i = 0; i < n; ++i) R.col_reord[i] = i;
        R.rp.swap(rp);
        R.ci.swap(ci);
        R.val.swap(val);
        R.entry_orig.swap(id);
        return R;
    }

    // (3) RCM or identity (sap/graph.h:560-575)
    std::vector<int> rcm_perm;
    if (o.doRCM) {
        R.k_reorder = rcm(n, rp, ci, R.col_reord, rcm_perm);
        if (R.k_reorder < 0) throw Error(SAP_ERR_HIP, "RCM failed");
        // permute rows and columns, regroup rows stably (sap/graph.h:1545-1577)
        std::vector<int> new_row(nnz);
        for (int i = 0; i < n; ++i)
            for (int j = rp[i]; j < rp[i + 1]; ++j) {
                new_row[j] = rcm_perm[i];
                ci[j] = rcm_perm[ci[j]];
            }
        regroup_rows(n, new_row, ci, val, rp, &id);
    } else {
        R.k_reorder = R.k_db;
        R.col_reord.resize(n);
        rcm_perm.resize(n);
        for (int i = 0; i < n; ++i) R.col_reord[i] = rcm_perm[i] = i;
    }

    R.time_reorder = since(t_reorder);

    // (4) drop-off (Graph::dropOff, sap/graph.h:600-747): entries bucketed by
    //     |row - col| (stable), whole outer diagonals removed while the kept
    //     sum of squares (in T) stays >= (1 - frac) of the total
    R.k = R.k_reorder;
    R.dropoff_actual = 0;
    if (R.k_reorder > o.maxBandwidth || o.dropOffFrac > 0) {
        const auto t_drop = clock::now();
        std::vector<int> bucket(n + 1, 0);
        for (int i = 0; i < n; ++i)
            for (int l = rp[i]; l < rp[i + 1]; ++l) bucket[std::abs(i - ci[l])]++;
        {
            int run = 0;
            for (int d = 0; d < n; ++d) { const int c = bucket[d]; bucket[d] = run; run += c; }
        }
        std::vector<int> cr(nnz), cc(nnz), cid(nnz);
        std::vector<T> cv(nnz);
        for (int i = 0; i < n; ++i)
            for (int l = rp[i]; l < rp[i + 1]; ++l) {
                const int d = bucket[std::abs(i - ci[l])]++;
                cr[d] = i;
                cc[d] = ci[l];
                cv[d] = val[l];
                cid[d] = id[l];
            }
        T norm_in = (T)0;
        for (size_t l = 0; l < nnz; ++l) norm_in += val[l] * val[l];
        const T frac = (T)o.dropOffFrac;
        const T min_norm_out = ((T)1 - frac) * norm_in;
        T norm_out = norm_in;
        // walk from the longest edges (the end of the bucketed array)
        // (an index past the front has no band: the walk stops there, where
        //  the reference would read outside its array)
        auto band_of = [&](long e) { return e < 0 ? -1 : std::abs(cr[e] - cc[e]); };
        long first = (long)nnz - 1, last = first;
        int final_k = band_of(first);
        {
            int bw;
            while ((bw = band_of(last)) > o.maxBandwidth) last--;
            T s = (T)0;
            for (long e = first; e > last; --e) s += cv[e] * cv[e];
            norm_out -= s;
            first = last;
            final_k = bw;
        }
        if (norm_out >= min_norm_out) {
            while (true) {
                const int bw = band_of(first);
                if (bw == 0) { final_k = 0; break; }
                do { last--; } while (band_of(last) == bw);
                T s = (T)0;
                for (long e = first; e > last; --e) s += cv[e] * cv[e];
                if (norm_out - s < min_norm_out) break;
                norm_out -= s;
                first = last;
                final_k = bw;
            }
        }
        R.dropoff_actual = (double)((T)1 - norm_out / norm_in);
        R.k = final_k;
        // keep entries [0, first] and regroup: by column (stable), then by row
        // filling each row from its end (so a row's columns end up descending)
        const size_t keep = (size_t)(first + 1);
        std::vector<int> cb(n + 1, 0);
        for (size_t e = 0; e < keep; ++e) cb[cc[e]]++;
        { int run = 0; for (int d = 0; d < n; ++d) { const int c = cb[d]; cb[d] = run; run += c; } }
        std::vector<int> r1(keep), c1(keep), i1(keep);
        std::vector<T> v1(keep);
        for (size_t e = 0; e < keep; ++e) {
            const int d = cb[cc[e]]++;
            r1[d] = cr[e];
            c1[d] = cc[e];
            v1[d] = cv[e];
            i1[d] = cid[e];
        }
        rp.assign(n + 1, 0);
        for (size_t e = 0; e < keep; ++e) rp[r1[e] + 1]++;
        for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
        std::vector<int> end(rp.begin() + 1, rp.end());
        ci.assign(keep, 0);
        val.assign(keep, (T)0);
        id.assign(keep, 0);
        for (size_t e = 0; e < keep; ++e) {
            const int d = --end[r1[e]];
            ci[d] = c1[e];
            val[d] = v1[e];
            id[d] = i1[e];
        }
        R.time_dropOff = since(t_drop);
    }

    // (5) final row permutation: DB then RCM (combinePermutation,
    //     sap/precond.h:1626-1636, :2094-2098)
    R.row_perm.resize(n);
    for (int i = 0; i < n; ++i) R.row_perm[i] = rcm_perm[R.db_perm[i]];
    R.rp.swap(rp);
    R.ci.swap(ci);
    R.val.swap(val);
    R.entry_orig.swap(id);
    return R;
}

namespace {

// partition of index l: the first n mod P partitions hold n/P + 1 rows
inline int part_of(int l, int part, int rem)
{
    int p = l / (part + 1);
    if (p >= rem) p = rem + (l - rem * (part + 1)) / part;
    return p;
}

// Per-partition RCM (Graph::partitionedRCM, sap/graph.h:1593-1757) over the
// diagonal-block graph: rows [nb, ne), their entries [eb, ee).
void partitioned_rcm(int n, const std::vector<int>& drp, const std::vector<int>& rows, const std::vector<int>& cols,
                     int eb, int ee, int nb, int ne, std::vector<int>& reord, std::vector<int>& perm,
                     std::vector<int>& buf)
{
    for (int i = nb; i < ne; ++i) reord[i] = i;
    int opt_bw = 0;
    for (int e = eb; e < ee; ++e) opt_bw = std::max(opt_bw, std::abs(rows[e] - cols[e]));
    // topology: both directions of every off-diagonal entry, grouped by
    // row (backward fill), offsets absolute by node
    std::vector<int> trp(n + 1, 0), tci;
    {
        std::vector<int> r2, c2;
        for (int e = eb; e < ee; ++e) {
            if (rows[e] == cols[e]) continue;
            r2.push_back(rows[e]); c2.push_back(cols[e]);
            r2.push_back(cols[e]); c2.push_back(rows[e]);
        }
        for (int r : r2) trp[r]++;
        int run = 0;
        for (int i = nb; i <= ne; ++i) { run += trp[i]; trp[i] = run; }
        tci.assign(r2.size(), 0);
        for (int e = (int)r2.size() - 1; e >= 0; --e) tci[--trp[r2[e]]] = c2[e];
    }
    const int MAX_NUM_TRIAL = 5, BANDWIDTH_THRESHOLD = 128;
    std::vector<char> tried(n, 0);
    // degrees of the partition's rows in the diagonal-block matrix; the
    // reference's transform covers ne - nb - 1 of them, so the last row's
    // degree stays 0 (sap/graph.h:1627)
    std::vector<int> pushed(ne, -1), levels(n, 0), degs(ne - nb, 0);
    for (int i = nb; i < ne - 1; ++i) degs[i - nb] = drp[i + 1] - drp[i];
    int max_level = 0, p_max_level = 0;   // the reference leaves them unset until the first push
    for (int trial = 0; trial < MAX_NUM_TRIAL; ++trial) {
        std::queue<int> q;
        std::priority_queue<NodeKey<int>, std::vector<NodeKey<int>>, KeyGreater<int>> pq;
        int node;
        if (trial > 0) {
            const int max_pos = (int)(std::max_element(levels.begin() + nb, levels.begin() + ne) - levels.begin());
            const int max_count = (int)std::count(levels.begin() + nb, levels.begin() + ne, max_level);
            if (max_count > 1) {
                std::vector<int> verts;
                for (int i = nb; i < ne; ++i)
                    if (levels[i] == max_level) verts.push_back(i);
                const int pos = (int)(std::min_element(degs.begin(), degs.begin() + max_count) - degs.begin());
                node = verts[pos];
            } else {
                node = max_pos;
            }
            while (tried[node]) node = (node + 1) % n;
        } else {
            node = (int)(std::min_element(degs.begin(), degs.end()) - degs.begin()) + nb;
        }
        tried[node] = 1;
        levels[node] = 0;
        pushed[node] = trial;
        q.push(node);
        int left = ne - nb, j = nb, last = nb;
        while (left--) {
            if (q.empty()) {
                left++;
                int i;
                for (i = last; i < ne; ++i) {
                    if (pushed[i] != trial) {
                        q.push(i);
                        pushed[i] = trial;
                        last = i;
                        break;
                    }
                }
                if (i < ne) continue;
                throw Error(SAP_ERR_HIP, "partitioned RCM ran out of vertices");
            }
            node = q.front();
            q.pop();
            buf[j++] = node;
            for (int e = trp[node]; e < trp[node + 1]; ++e) {
                const int t = tci[e];
                if (pushed[t] != trial) {
                    pushed[t] = trial;
                    pq.push({t, trp[t + 1] - trp[t]});
                    max_level = levels[t] = levels[node] + 1;
                }
            }
            while (!pq.empty()) {
                q.push(pq.top().node);
                pq.pop();
            }
        }
        for (int i = nb; i < ne; ++i) perm[buf[i]] = i;
        int bw = 0;
        for (int e = eb; e < ee; ++e) bw = std::max(bw, std::abs(perm[rows[e]] - perm[cols[e]]));
        if (opt_bw > bw) {
            opt_bw = bw;
            std::copy(buf.begin() + nb, buf.begin() + ne, reord.begin() + nb);
        }
        if (trial > 0) {
            if (p_max_level >= max_level) break;
            const double ratio = 1.0 * (max_level - p_max_level) / p_max_level;
            if (ratio < 0.01) break;
            p_max_level = max_level;
        } else {
            p_max_level = max_level;
        }
        if (opt_bw <= BANDWIDTH_THRESHOLD) break;
    }
    for (int i = nb; i < ne; ++i) perm[reord[i]] = i;
}

}  // namespace

template <class T>
VarBand<T> assemble_var(int n, int k, int P, const std::vector<int>& rp, const std::vector<int>& ci,
                        const std::vector<T>& val)
{
    VarBand<T> V;
    V.n = n;
    V.k = k;
    V.P = P;
    const int part = n / P, rem = n % P;
    const size_t kk = (size_t)k * k;

    // (1) Graph::assembleOffDiagMatrices (sap/graph.h:756-898)
    V.WV.assign(2 * kk * (P - 1), (T)0);
    V.offDiags.assign(2 * kk * (P - 1), (T)0);
    V.widths_left.assign(P - 1, 0);
    V.widths_right.assign(P - 1, 0);
    V.perms_left.assign((size_t)(P - 1) * k, -1);
    V.perms_right.assign((size_t)(P - 1) * k, -1);
    std::vector<int> reord_left((size_t)(P - 1) * k, -1), reord_right((size_t)(P - 1) * k, -1);
    V.exists.assign(n, 0);
    std::vector<int> drows;
    V.dci.clear();
    V.dval.clear();
    for (int j = 0; j < n; ++j) {
        for (int e = rp[j]; e < rp[j + 1]; ++e) {
            const int l = ci[e];
            const int cp = part_of(l, part, rem), rpn = part_of(j, part, rem);
            if (cp == rpn) {
                drows.push_back(j);
                V.dci.push_back(l);
                V.dval.push_back(val[e]);
            } else if (cp > rpn) {   // V / B block: row in partition rpn, column in cp
                V.exists[j] = 1;
                const int end_row = part * rpn + (rpn < rem ? rpn + part + 1 : rem + part);
                const int start_col = part * cp + (cp < rem ? cp : rem);
                const size_t slot = (size_t)rpn * k + (l - start_col);
                if (reord_right[slot] < 0) {
                    reord_right[slot] = V.widths_right[rpn];
                    V.perms_right[(size_t)rpn * k + V.widths_right[rpn]] = l - start_col;
                    V.widths_right[rpn]++;
                }
                const size_t base = (size_t)rpn * 2 * kk;
                V.offDiags[base + (size_t)(j + k - end_row) * k + (l - start_col)] = val[e];
                V.WV[base + (size_t)(j + k - end_row) + (size_t)reord_right[slot] * k] = val[e];
            } else {                 // W / C block: row in partition rpn, column in cp < rpn
                const int start_row = part * rpn + (rpn < rem ? rpn : rem);
                const int end_col = part * cp + (cp < rem ? cp + part + 1 : rem + part);
                const size_t slot = (size_t)(rpn - 1) * k + (l - end_col + k);
                if (reord_left[slot] < 0) {
                    reord_left[slot] = k - 1 - V.widths_left[rpn - 1];
                    V.perms_left[(size_t)(rpn - 1) * k + k - 1 - V.widths_left[rpn - 1]] = l - end_col + k;
                    V.widths_left[rpn - 1]++;
                }
                const size_t base = (size_t)(cp * 2 + 1) * kk;
                V.offDiags[base + (size_t)(j - start_row) * k + (l - end_col + k)] = val[e];
                V.WV[base + (size_t)(j - start_row) + (size_t)reord_left[slot] * k] = val[e];
            }
        }
    }
    V.drp.assign(n + 1, 0);
    for (int r : drows) V.drp[r + 1]++;
    for (int i = 0; i < n; ++i) V.drp[i + 1] += V.drp[i];

    // (2) secondLevelReordering (sap/graph.h:908-1008)
    V.second_reorder.assign(n, 0);
    V.second_perm.assign(n, 0);
    {
        std::vector<int> buf(n, 0);
        int nb = 0, eb = 0;
        for (int i = 0; i < P; ++i) {
            const int ne = nb + part + (i < rem ? 1 : 0);
            const int ee = V.drp[ne];
            partitioned_rcm(n, V.drp, drows, V.dci, eb, ee, nb, ne, V.second_reorder, V.second_perm, buf);
            nb = ne;
            eb = ee;
        }
        // permute the diagonal blocks and regroup their rows (backward fill)
        const size_t dn = V.dci.size();
        std::vector<int> nr(dn), nc(dn);
        for (size_t e = 0; e < dn; ++e) {
            nr[e] = V.second_perm[drows[e]];
            nc[e] = V.second_perm[V.dci[e]];
        }
        regroup_rows(n, nr, nc, V.dval, V.drp);
        V.dci.swap(nc);
    }
    V.first_rows.assign(P - 1, 0);
    {
        int last_row = 0;
        for (int i = 0; i < P - 1; ++i) {
            last_row += part + (i < rem ? 1 : 0);
            V.first_rows[i] = last_row;
            for (int j = last_row - k; j < last_row; ++j)
                if (j >= 0 && V.exists[j] && V.first_rows[i] > V.second_perm[j]) V.first_rows[i] = V.second_perm[j];
        }
    }

    // (3) assembleBandedMatrix, several partitions (sap/graph.h:1093-1206)
    V.ks.assign(P, 0);
    for (int j = 0; j < n; ++j)
        for (int e = V.drp[j]; e < V.drp[j + 1]; ++e) {
            const int l = V.dci[e];
            const int cp = part_of(l, part, rem);
            V.ks[cp] = std::max(V.ks[cp], std::abs(l - j));
        }
    V.boffsets.assign(P + 1, 0);
    for (int i = 0; i < P; ++i)
        V.boffsets[i + 1] = V.boffsets[i] + (int64_t)(part + (i < rem ? 1 : 0)) * (2 * V.ks[i] + 1);
    V.ks_col.assign(n, 0);
    V.ks_row.assign(n, 0);
    for (int j = 0; j < n; ++j)
        for (int e = V.drp[j]; e < V.drp[j + 1]; ++e) {
            const int l = V.dci[e];
            V.ks_col[l] = std::max(V.ks_col[l], j - l);
            V.ks_row[j] = std::max(V.ks_row[j], l - j);
        }
    {
        int begin = 0;
        for (int i = 0; i < P; ++i) {
            const int end = begin + part + (i < rem ? 1 : 0);
            for (int j = begin + 1; j < end; ++j) {
                V.ks_col[j] = std::min(std::max(V.ks_col[j], V.ks_col[j - 1] - 1), end - j - 1);
                V.ks_row[j] = std::min(std::max(V.ks_row[j], V.ks_row[j - 1] - 1), end - j - 1);
            }
            begin = end;
        }
    }
    return V;
}

template VarBand<float> assemble_var<float>(int, int, int, const std::vector<int>&, const std::vector<int>&,
                                            const std::vector<float>&);
template VarBand<double> assemble_var<double>(int, int, int, const std::vector<int>&, const std::vector<int>&,
                                              const std::vector<double>&);

BandProfile band_profile(int n, const int* rp, const int* ci, int P)
{
    std::vector<int> ks_row(n, 0), ks_col(n, 0);
    const int part = n / P, rem = n % P;
    auto part_of = [&](int l) {
        int p = l / (part + 1);
        if (p >= rem) p = rem + (l - rem * (part + 1)) / part;
        return p;
    };
    for (int j = 0; j < n; ++j)
        for (int e = rp[j]; e < rp[j + 1]; ++e) {
            const int l = ci[e];
            if (P > 1 && part_of(l) != part_of(j)) continue;   // coupling entries
            ks_col[l] = std::max(ks_col[l], j - l);
            ks_row[j] = std::max(ks_row[j], l - j);
        }
    // extents may shrink by at most one per row / column; with several
    // partitions they are also clipped to the partition's end (the first
    // row of each partition is left as it is)
    if (P == 1) {
        for (int i = 1; i < n; ++i) {
            ks_col[i] = std::max(ks_col[i], ks_col[i - 1] - 1);
            ks_row[i] = std::max(ks_row[i], ks_row[i - 1] - 1);
        }
    } else {
        int begin = 0;
        for (int p = 0; p < P; ++p) {
            const int end = begin + part + (p < rem ? 1 : 0);
            for (int j = begin + 1; j < end; ++j) {
                ks_col[j] = std::min(std::max(ks_col[j], ks_col[j - 1] - 1), end - j - 1);
                ks_row[j] = std::min(std::max(ks_row[j], ks_row[j - 1] - 1), end - j - 1);
            }
            begin = end;
        }
    }
    BandProfile b;
    for (int i = 0; i < n; ++i) {
        b.sum_row += ks_row[i];
        b.sum_col += ks_col[i];
        b.sum_prod += (double)ks_row[i] * ks_col[i];
    }
    return b;
}

template Reordering<float> reorder_csr<float>(int, const int*, const int*, const double*, const GraphOptions&);
template Reordering<double> reorder_csr<double>(int, const int*, const int*, const double*, const GraphOptions&);

}  // namespace sap_amd
