"""Plain-Python restatement of the reference's host reordering pipeline for
CSR input: DB (diagonal boosting), RCM and drop-off, and the permutation
composition of transformToBandedMatrix.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the shipped host code
(saplibrary_amd/csrc/graph.hip through the C ABI).  Nothing in the shipped
library imports it.  Pure-Python loops: small matrices only (n up to a few
thousand).

Follows spikegpu/SaPLibrary
  sap/precond.h:1656-1713   values to the preconditioner type, zeros removed
  sap/graph.h:446-579       Graph::reorder
  sap/graph.h:1214-1361     Graph::DB
  sap/graph.h:1827-1857     get_csc_matrix (+ sap/device/db.cuh:12-37)
  sap/graph.h:1875-2035     init_reduced_cval
  sap/graph.h:2043-2173     find_shortest_aug_path
  sap/graph.h:1371-1584     Graph::RCM, :1764-1809 buildTopology
  sap/graph.h:598-747       Graph::dropOff
  sap/precond.h:1626-1636   combinePermutation
  sap/graph.h:1019-1206     assembleBandedMatrix (band extents ks_row/ks_col)

Every priority queue is std::priority_queue over a std::vector with the
reference's CompareValue (a.key > b.key); LibstdcxxPQ reproduces libstdc++'s
push_heap / pop_heap so that equal keys leave the queue in the same order.

Arithmetic precision follows the reference: DB costs and duals in double
(math.log / math.exp are the C library's log / exp, as std::log / std::exp),
matrix values, scalings and drop-off sums in the preconditioner type T
(numpy.float32 or float64 scalars, summed left to right).
"""
import math

import numpy as np

LOC_INFINITY = 1e37           # Graph<T>::LOC_INFINITY (sap/graph.h:417)


class LibstdcxxPQ:
    """std::priority_queue<(node, key), vector, CompareValue> where
    CompareValue(a, b) = a.key > b.key (sap/graph.h:387-391): the top is a
    smallest key.  push/pop follow libstdc++'s bits/stl_heap.h
    (__push_heap, __adjust_heap, __pop_heap)."""

    def __init__(self):
        self.c = []

    @staticmethod
    def _comp(a, b):
        return a[1] > b[1]

    def _push_heap(self, hole, top, value):
        c = self.c
        parent = (hole - 1) // 2
        while hole > top and self._comp(c[parent], value):
            c[hole] = c[parent]
            hole = parent
            parent = (hole - 1) // 2
        c[hole] = value

    def _adjust_heap(self, hole, length, value):
        c = self.c
        top = hole
        second = hole
        while second < (length - 1) // 2:
            second = 2 * (second + 1)
            if self._comp(c[second], c[second - 1]):
                second -= 1
            c[hole] = c[second]
            hole = second
        if (length & 1) == 0 and second == (length - 2) // 2:
            second = 2 * (second + 1)
            c[hole] = c[second - 1]
            hole = second - 1
        self._push_heap(hole, top, value)

    def push(self, node, key):
        self.c.append((node, key))
        self._push_heap(len(self.c) - 1, 0, self.c[-1])

    def top(self):
        return self.c[0]

    def pop(self):
        c = self.c
        if len(c) > 1:
            last = len(c) - 1
            value = c[last]
            c[last] = c[0]
            self._adjust_heap(0, last, value)
        c.pop()

    def empty(self):
        return not self.c


def _T(prec):
    return np.float32 if prec == "float" else np.float64


def _regroup_rows_backward(n, new_row, cols, vals):
    """Counting sort by row, filling each row from its end while walking the
    entries backwards (sap/graph.h:1334-1347, :1545-1577): stable."""
    nnz = len(new_row)
    off = [0] * (n + 1)
    for r in new_row:
        off[r] += 1
    run = 0
    for i in range(n + 1):
        run += off[i]
        off[i] = run                        # inclusive scan
    ci = [0] * nnz
    va = [None] * nnz
    for e in range(nnz - 1, -1, -1):
        off[new_row[e]] -= 1
        d = off[new_row[e]]
        ci[d] = cols[e]
        va[d] = vals[e]
    return off, ci, va                      # off[] now holds the row starts


# ------------------------------------------------------------------------ DB
def _reduce_by_key_rows(rp, vals, op):
    """thrust::reduce_by_key over the row index of each entry: one value per
    run of equal keys, i.e. per non-empty row, stored compactly from 0."""
    n = len(rp) - 1
    out = []
    for i in range(n):
        if rp[i] == rp[i + 1]:
            continue
        m = vals[rp[i]]
        for j in range(rp[i] + 1, rp[i + 1]):
            m = op(m, vals[j])
        out.append(m)
    return out


def _db(n, rp, ci, val, first_stage_only):
    """Graph::DB up to the scalings (sap/graph.h:1214-1285): returns
    (match[col] = row, u, v, rowmax)."""
    nnz = len(ci)
    # get_csc_matrix + getResidualValues: c = log(max / |a|), the maximum
    # read at index `row` from the compacted per-row maxima
    c = [abs(float(val[j])) for j in range(nnz)]
    mx = _reduce_by_key_rows(rp, c, max)
    mx += [0.0] * (n - len(mx))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            c[j] = math.log(mx[i] / c[j])

    # init_reduced_cval (sap/graph.h:1875-2035)
    u = _reduce_by_key_rows(rp, c, min)      # per non-empty row, used by column
    u += [0.0] * (n - len(u))
    v = [LOC_INFINITY] * n
    match = [0] * n
    rev = [0] * nnz if nnz > n else [0] * max(nnz, n)
    matched = [False] * n
    rev_matched = [False] * n
    for i in range(n):
        min_idx = -1
        for j in range(rp[i], rp[i + 1]):
            if c[j] > LOC_INFINITY / 2.0:
                continue
            t = c[j] - u[ci[j]]
            if v[i] > t:
                v[i] = t
                min_idx = j
        if min_idx >= 0:
            col = ci[min_idx]
            if not matched[col]:
                rev_matched[i] = True
                matched[col] = True
                match[col] = i
                rev[i] = min_idx
    if first_stage_only:
        for i in range(n):
            if rev_matched[i]:
                continue
            cur_v = int(v[i])                 # int cur_v_val = v_val[i]
            found = False
            for j in range(rp[i], rp[i + 1]):
                if c[j] > LOC_INFINITY / 2.0:
                    continue
                col = ci[j]
                if not matched[col]:
                    continue
                if c[j] - u[col] - cur_v > 1e-10:
                    continue
                i2 = match[col]
                for j2 in range(rp[i2], rp[i2 + 1]):
                    if c[j2] > LOC_INFINITY / 2.0:
                        continue
                    col2 = ci[j2]
                    if matched[col2]:
                        continue
                    matched[col2] = True
                    match[col2] = i2
                    rev[i2] = j2
                    match[col] = i
                    rev_matched[i] = True
                    rev[i] = j
                    found = True
                    break
                if found:
                    break
        for i in range(n):
            if rev_matched[i]:
                continue
            for j in range(rp[i], rp[i + 1]):
                col = ci[j]
                if not matched[col]:
                    matched[col] = True
                    rev_matched[i] = True
                    match[col] = i
                    rev[i] = j
                    break
        for i in range(n):
            if rev_matched[i]:
                continue
            found = False
            for j in range(rp[i], rp[i + 1]):
                col = ci[j]
                i2 = match[col]
                for j2 in range(rp[i2], rp[i2 + 1]):
                    col2 = ci[j2]
                    if matched[col2]:
                        continue
                    found = True
                    matched[col2] = True
                    match[col2] = i2
                    rev[i2] = j2
                    match[col] = i
                    rev_matched[i] = True
                    rev[i] = j
                    break
                if found:
                    break
        last_j = 0
        for i in range(n):
            if rev_matched[i]:
                continue
            for j in range(last_j, n):
                if matched[j]:
                    continue
                rev_matched[i] = matched[j] = True
                match[j] = i
                last_j = j + 1
                break
        u = [1.0] * n
        v = [1.0] * n
    else:
        for i in range(n):
            if not matched[i]:
                u[i] = 0.0
            if not rev_matched[i]:
                v[i] = 0.0

    # find_shortest_aug_path for every unmatched row (sap/graph.h:2043-2173);
    # the Dijkstra scratch (inB, visited, d, B) lives across calls
    inB = [False] * n
    visited = [False] * n
    dval = [LOC_INFINITY] * n
    Blist = [0] * n
    prev = [0] * n
    irn = [0] * n
    for init in range(n):
        if rev_matched[init]:
            continue
        Q = LibstdcxxPQ()
        lsp, lsap = 0.0, LOC_INFINITY
        cur = init
        isap = ksap = -1
        b_cnt = 0
        prev[init] = -1
        while True:
            for e in range(rp[cur], rp[cur + 1]):
                col = ci[e]
                if inB[col]:
                    continue
                if c[e] > LOC_INFINITY / 2.0:
                    continue
                reduced = c[e] - u[col] - v[cur]
                if reduced + 1e-10 < 0:
                    raise ValueError("Negative_DB_weight")
                d_new = lsp + reduced
                if d_new < lsap:
                    if not matched[col]:
                        lsap = d_new
                        isap = col
                        ksap = e
                        match[isap] = cur
                    elif d_new < dval[col]:
                        dval[col] = d_new
                        prev[match[col]] = cur
                        Q.push(col, d_new)
                        irn[col] = e
            found = False
            top = None
            while not Q.empty():
                top = Q.top()
                Q.pop()
                if visited[top[0]]:
                    continue
                found = True
                break
            if not found:
                break
            t = top[0]
            visited[t] = True
            lsp = top[1]
            if lsap <= lsp:
                visited[t] = False
                dval[t] = LOC_INFINITY
                break
            inB[t] = True
            Blist[b_cnt] = t
            b_cnt += 1
            cur = match[t]
        if lsap < LOC_INFINITY / 2.0:
            matched[isap] = True
            cur = match[isap]
            v[cur] = c[ksap]
            while prev[cur] >= 0:
                match[isap] = cur
                nk = rev[cur]
                ni = ci[nk]
                nk = irn[ni]
                rev[cur] = ksap
                cur = prev[cur]
                isap = ni
                ksap = nk
            match[isap] = cur
            rev[cur] = ksap
            rev_matched[cur] = True
            for b in range(b_cnt):
                row = Blist[b]
                jv = match[row]
                tk = rev[jv]
                u[row] += dval[row] - lsap
                v[jv] = c[tk] - u[row]
                dval[row] = LOC_INFINITY
                visited[row] = False
                inB[row] = False
            while not Q.empty():
                dval[Q.top()[0]] = LOC_INFINITY
                Q.pop()
    if not all(matched):
        raise ValueError("Matrix_singular")
    return match, u, v, mx


# ----------------------------------------------------------------------- RCM
def _rcm(n, rp, ci):
    """Graph::RCM (sap/graph.h:1371-1584).  Returns (bandwidth, reord, perm)
    with reord[new] = old and perm[old] = new."""
    nnz = len(ci)
    rows = [0] * nnz
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            rows[j] = i
    deg = [rp[i + 1] - rp[i] for i in range(n)]
    bandwidth = max((abs(rows[e] - ci[e]) for e in range(nnz)), default=0)
    # buildTopology: both directions of every off-diagonal entry, in entry
    # order, grouped by row with the backward fill
    r2, c2 = [], []
    for e in range(nnz):
        a, b = rows[e], ci[e]
        if a != b:
            r2 += [a, b]
            c2 += [b, a]
    arp, aci, _ = _regroup_rows_backward(n, r2, c2, [0] * len(r2))

    reord = list(range(n))
    tried = [False] * n
    pushed = [-1] * n
    levels = [0] * n
    max_level = 0
    p_max_level = 0
    for trial in range(5):                        # MAX_NUM_TRIAL
        q = []
        qh = 0
        pq = LibstdcxxPQ()
        if trial > 0:
            max_pos = levels.index(max(levels))
            max_count = levels.count(max_level)
            if max_count > 1:
                verts = [i for i in range(n) if levels[i] == max_level]
                val = deg[:max_count]             # gather of ori_degrees[0 .. max_count-1]
                node = verts[val.index(min(val))]
            else:
                node = max_pos
            while tried[node]:
                node = (node + 1) % n
        else:
            node = deg.index(min(deg))
        tried[node] = True
        levels[node] = 0
        pushed[node] = trial
        q.append(node)
        tmp = []
        left, last = n, 0
        while left > 0:
            left -= 1
            if qh == len(q):
                left += 1
                i = last
                while i < n:
                    if pushed[i] != trial:
                        q.append(i)
                        pushed[i] = trial
                        last = i
                        break
                    i += 1
                if i < n:
                    continue
                raise RuntimeError("RCM: no vertex left")
            node = q[qh]
            qh += 1
            tmp.append(node)
            lev = levels[node]
            for e in range(arp[node], arp[node + 1]):
                t = aci[e]
                if pushed[t] != trial:
                    pushed[t] = trial
                    pq.push(t, deg[t])
                    levels[t] = lev + 1
            while not pq.empty():
                q.append(pq.top()[0])
                pq.pop()
        tperm = [0] * n
        for i, node in enumerate(tmp):
            tperm[node] = i
        bw = max((abs(tperm[rows[e]] - tperm[ci[e]]) for e in range(nnz)), default=0)
        if bandwidth > bw:
            bandwidth = bw
            reord = tmp
        if trial > 0:
            if p_max_level >= max_level:
                break
            if 1.0 * (max_level - p_max_level) / p_max_level < 0.01:
                break
            p_max_level = max_level
        else:
            p_max_level = max_level
        if bandwidth <= 256:                      # BANDWIDTH_THRESHOLD
            break
    perm = [0] * n
    for i, node in enumerate(reord):
        perm[node] = i
    return bandwidth, reord, perm


# ------------------------------------------------------------------ drop-off
def _drop_off(n, rp, ci, val, frac, max_bw, T):
    """Graph::dropOff (sap/graph.h:598-747): returns (k, frac_actual, rp, ci, val)."""
    nnz = len(ci)
    bucket = [0] * n
    for i in range(n):
        for l in range(rp[i], rp[i + 1]):
            bucket[abs(i - ci[l])] += 1
    run = 0
    for d in range(n):
        cnt = bucket[d]
        bucket[d] = run
        run += cnt
    cr, cc, cv = [0] * nnz, [0] * nnz, [None] * nnz
    for i in range(n):
        for l in range(rp[i], rp[i + 1]):
            d = bucket[abs(i - ci[l])]
            bucket[abs(i - ci[l])] += 1
            cr[d], cc[d], cv[d] = i, ci[l], val[l]
    norm_in = T(0)
    for l in range(nnz):
        norm_in = T(norm_in + val[l] * val[l])
    min_norm_out = T((T(1) - T(frac)) * norm_in)
    norm_out = norm_in

    def band(e):
        return abs(cr[e] - cc[e]) if e >= 0 else -1

    first = last = nnz - 1
    bw = band(last)
    while bw > max_bw:
        last -= 1
        bw = band(last)
    s = T(0)
    for e in range(first, last, -1):
        s = T(s + cv[e] * cv[e])
    norm_out = T(norm_out - s)
    first = last
    final_k = bw
    if norm_out >= min_norm_out:
        while True:
            bw = band(first)
            if bw == 0:
                final_k = 0
                break
            last -= 1
            while band(last) == bw:
                last -= 1
            s = T(0)
            for e in range(first, last, -1):
                s = T(s + cv[e] * cv[e])
            if T(norm_out - s) < min_norm_out:
                break
            norm_out = T(norm_out - s)
            first = last
            final_k = bw
    frac_actual = T(T(1) - T(norm_out / norm_in))
    keep = first + 1
    # regroup: by column (forward, stable), then rows filled by decrementing
    # their inclusive-scan ends in forward order (columns end up descending)
    cb = [0] * n
    for e in range(keep):
        cb[cc[e]] += 1
    run = 0
    for d in range(n):
        cnt = cb[d]
        cb[d] = run
        run += cnt
    r1, c1, v1 = [0] * keep, [0] * keep, [None] * keep
    for e in range(keep):
        d = cb[cc[e]]
        cb[cc[e]] += 1
        r1[d], c1[d], v1[d] = cr[e], cc[e], cv[e]
    rb = [0] * n
    for e in range(keep):
        rb[r1[e]] += 1
    run = 0
    for d in range(n):
        run += rb[d]
        rb[d] = run
    ci2, va2 = [0] * keep, [None] * keep
    for e in range(keep):
        rb[r1[e]] -= 1
        d = rb[r1[e]]
        ci2[d], va2[d] = c1[e], v1[e]
    rp2 = [0] * (n + 1)
    for e in range(keep):
        rp2[r1[e] + 1] += 1
    for i in range(n):
        rp2[i + 1] += rp2[i]
    return final_k, float(frac_actual), rp2, ci2, va2


# ------------------------------------------------------------------ pipeline
def reorder(rp, ci, val, do_db=True, scale=True, first_stage_only=False, max_bw=2 ** 31 - 1, drop_off=0.0,
            prec="float"):
    """transformToBandedMatrix's host part (sap/precond.h:1656-1776 with
    Graph::reorder and dropOff).  Inputs: CSR (row offsets, column indices,
    FP64 values).  Returns a dict with the same fields as
    saplibrary_amd.reorder_csr.  Raises ValueError("Matrix_singular" /
    "Negative_DB_weight")."""
    T = _T(prec)
    rp = [int(x) for x in rp]
    n = len(rp) - 1
    # (1) values in T, zeros removed (sap/precond.h:1694-1711)
    nrp, nci, nval = [0] * (n + 1), [], []
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            x = T(val[j])
            if x != 0:
                nci.append(int(ci[j]))
                nval.append(x)
        nrp[i + 1] = len(nci)
    rp, ci, val = nrp, nci, nval
    nnz = len(ci)

    # (2) DB (Graph::reorder, sap/graph.h:494-520; Graph::DB :1214-1361)
    row_scale = [T(1)] * n
    col_scale = [T(1)] * n
    rs = cs = [1.0] * n
    db_perm = list(range(n))
    if do_db:
        match, u, v, mx = _db(n, rp, ci, val, first_stage_only)
        cs = [math.exp(u[i]) for i in range(n)]
        rs = [math.exp(v[i]) / mx[i] for i in range(n)]
        db_perm = [0] * n
        for i in range(n):
            db_perm[match[i]] = i
        do_scale = scale and not first_stage_only
        new_row = [0] * nnz
        vals2 = list(val)
        for i in range(n):
            for l in range(rp[i], rp[i + 1]):
                new_row[l] = db_perm[i]
                if do_scale:
                    vals2[l] = T(T(rs[i] * cs[ci[l]]) * val[l])
        row_scale = [T(x) for x in rs]
        col_scale = [T(x) for x in cs]
        rp, ci, val = _regroup_rows_backward(n, new_row, ci, vals2)
    k_db = 0
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            k_db = max(k_db, abs(i - ci[j]))

    # (3) RCM when maxBandwidth > 64 (sap/precond.h:1737-1738)
    if max_bw > 64:
        k_reorder, col_reord, rcm_perm = _rcm(n, rp, ci)
        new_row = [0] * nnz
        newc = [0] * nnz
        for i in range(n):
            for j in range(rp[i], rp[i + 1]):
                new_row[j] = rcm_perm[i]
                newc[j] = rcm_perm[ci[j]]
        rp, ci, val = _regroup_rows_backward(n, new_row, newc, val)
    else:
        k_reorder = k_db
        col_reord = list(range(n))
        rcm_perm = list(range(n))

    # (4) drop-off (sap/precond.h:1760-1776)
    k, dropoff = k_reorder, 0.0
    if k_reorder > max_bw or drop_off > 0:
        k, dropoff, rp, ci, val = _drop_off(n, rp, ci, val, drop_off, max_bw, T)

    # (5) row permutation: DB then RCM (combinePermutation, sap/precond.h:2094-2098)
    row_perm = [rcm_perm[db_perm[i]] for i in range(n)]
    return dict(
        k=k, k_db=k_db, k_reorder=k_reorder, dropoff=dropoff,
        db_perm=np.array(db_perm, dtype=np.int32), row_perm=np.array(row_perm, dtype=np.int32),
        col_reord=np.array(col_reord, dtype=np.int32),
        row_scale=np.array(row_scale, dtype=T).astype(np.float64),
        col_scale=np.array(col_scale, dtype=T).astype(np.float64),
        # the scalings before rounding to T (an entry's scale factor is
        # T(row_scale_d[i] * col_scale_d[j]), sap/graph.h:1307)
        row_scale_d=np.array(rs, dtype=np.float64), col_scale_d=np.array(cs, dtype=np.float64),
        rp=np.array(rp, dtype=np.int32), ci=np.array(ci, dtype=np.int32),
        val=np.array(val, dtype=T).astype(np.float64),
    )


def band_profile(n, rp, ci, P):
    """Sums of the band extents ks_row, ks_col of assembleBandedMatrix
    (sap/graph.h:1019-1091 for P == 1, :1093-1206 over the diagonal blocks
    for P > 1): (sum ks_row, sum ks_col, sum ks_row * ks_col)."""
    part, rem = n // P, n % P

    def part_of(l):
        p = l // (part + 1)
        if p >= rem:
            p = rem + (l - rem * (part + 1)) // part
        return p

    ks_row, ks_col = [0] * n, [0] * n
    for j in range(n):
        for e in range(rp[j], rp[j + 1]):
            l = int(ci[e])
            if P > 1 and part_of(l) != part_of(j):
                continue
            ks_col[l] = max(ks_col[l], j - l)
            ks_row[j] = max(ks_row[j], l - j)
    if P == 1:
        for i in range(1, n):
            ks_col[i] = max(ks_col[i], ks_col[i - 1] - 1)
            ks_row[i] = max(ks_row[i], ks_row[i - 1] - 1)
    else:
        begin = 0
        for p in range(P):
            end = begin + part + (1 if p < rem else 0)
            for j in range(begin + 1, end):
                ks_col[j] = min(max(ks_col[j], ks_col[j - 1] - 1), end - j - 1)
                ks_row[j] = min(max(ks_row[j], ks_row[j - 1] - 1), end - j - 1)
            begin = end
    return float(sum(ks_row)), float(sum(ks_col)), float(sum(a * b for a, b in zip(ks_row, ks_col)))


def csr_to_band(n, k, rp, ci, val):
    """Row-major FP64 band (n x (2k+1)) of a CSR matrix whose entries all lie
    within the half-bandwidth k (copyFromCOOMatrixToBandedMatrix,
    sap/device/data_transfer.cuh:421-448, in the row-major input layout)."""
    band = np.zeros((n, 2 * k + 1))
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            band[i, k + int(ci[e]) - i] = val[e]
    return band


# ----------------------------------------------------- variable-bandwidth assembly
def _part_of(l, part_size, rem):
    p = l // (part_size + 1)
    if p >= rem:
        p = rem + (l - rem * (part_size + 1)) // part_size
    return p


def _partitioned_rcm(n, drp, rows, cols, eb, ee, nb, ne, reord, perm, buf):
    """Graph::partitionedRCM (sap/graph.h:1593-1757) on rows [nb, ne) of the
    diagonal-block matrix (entries [eb, ee))."""
    for i in range(nb, ne):
        reord[i] = i
    opt_bw = max((abs(rows[e] - cols[e]) for e in range(eb, ee)), default=0)
    # buildTopology over this partition's entries
    r2, c2 = [], []
    for e in range(eb, ee):
        if rows[e] != cols[e]:
            r2 += [rows[e], cols[e]]
            c2 += [cols[e], rows[e]]
    trp = [0] * (n + 1)
    for r in r2:
        trp[r] += 1
    run = 0
    for i in range(nb, ne + 1):
        run += trp[i]
        trp[i] = run
    tci = [0] * len(r2)
    for e in range(len(r2) - 1, -1, -1):
        trp[r2[e]] -= 1
        tci[trp[r2[e]]] = c2[e]
    tried = [False] * n
    pushed = [-1] * ne
    levels = [0] * n
    # ori_degrees: the transform covers ne - nb - 1 rows (the last stays 0)
    degs = [0] * (ne - nb)
    for i in range(nb, ne - 1):
        degs[i - nb] = drp[i + 1] - drp[i]
    max_level = p_max_level = 0
    for trial in range(5):
        q, qh = [], 0
        pq = LibstdcxxPQ()
        if trial > 0:
            seg = levels[nb:ne]
            max_pos = nb + seg.index(max(seg))
            max_count = seg.count(max_level)
            if max_count > 1:
                verts = [i for i in range(nb, ne) if levels[i] == max_level]
                val = degs[:max_count]
                node = verts[val.index(min(val))]
            else:
                node = max_pos
            while tried[node]:
                node = (node + 1) % n
        else:
            node = degs.index(min(degs)) + nb
        tried[node] = True
        levels[node] = 0
        pushed[node] = trial
        q.append(node)
        left, j, last = ne - nb, nb, nb
        while left > 0:
            left -= 1
            if qh == len(q):
                left += 1
                i = last
                while i < ne:
                    if pushed[i] != trial:
                        q.append(i)
                        pushed[i] = trial
                        last = i
                        break
                    i += 1
                if i < ne:
                    continue
                raise RuntimeError("partitioned RCM: no vertex left")
            node = q[qh]
            qh += 1
            buf[j] = node
            j += 1
            for e in range(trp[node], trp[node + 1]):
                t = tci[e]
                if pushed[t] != trial:
                    pushed[t] = trial
                    pq.push(t, trp[t + 1] - trp[t])
                    levels[t] = levels[node] + 1
                    max_level = levels[t]
            while not pq.empty():
                q.append(pq.top()[0])
                pq.pop()
        for i in range(nb, ne):
            perm[buf[i]] = i
        bw = max((abs(perm[rows[e]] - perm[cols[e]]) for e in range(eb, ee)), default=0)
        if opt_bw > bw:
            opt_bw = bw
            reord[nb:ne] = buf[nb:ne]
        if trial > 0:
            if p_max_level >= max_level:
                break
            if 1.0 * (max_level - p_max_level) / p_max_level < 0.01:
                break
            p_max_level = max_level
        else:
            p_max_level = max_level
        if opt_bw <= 128:
            break
    for i in range(nb, ne):
        perm[reord[i]] = i


def assemble_var(n, k, P, rp, ci, val, prec="float"):
    """The variable-bandwidth assembly of transformToBandedMatrix
    (sap/precond.h:1785-1800): Graph::assembleOffDiagMatrices
    (sap/graph.h:756-898), secondLevelReordering (:908-1008) with
    partitionedRCM, assembleBandedMatrix for several partitions
    (:1093-1206).  rp/ci/val: the reordered matrix (values already in T);
    P: the clamped partition count (> 1)."""
    T = _T(prec)
    part, rem = n // P, n % P
    kk = k * k
    WV = [T(0)] * (2 * kk * (P - 1))
    off = [T(0)] * (2 * kk * (P - 1))
    wl, wr = [0] * (P - 1), [0] * (P - 1)
    pl, pr = [-1] * ((P - 1) * k), [-1] * ((P - 1) * k)
    rl, rr = [-1] * ((P - 1) * k), [-1] * ((P - 1) * k)
    exists = [False] * n
    drows, dcols, dvals = [], [], []
    for j in range(n):
        for e in range(rp[j], rp[j + 1]):
            l = int(ci[e])
            cp, rpn = _part_of(l, part, rem), _part_of(j, part, rem)
            v = T(val[e])
            if cp == rpn:
                drows.append(j)
                dcols.append(l)
                dvals.append(v)
            elif cp > rpn:
                exists[j] = True
                end_row = part * rpn + (rpn + part + 1 if rpn < rem else rem + part)
                start_col = part * cp + (cp if cp < rem else rem)
                s = rpn * k + l - start_col
                if rr[s] < 0:
                    rr[s] = wr[rpn]
                    pr[rpn * k + wr[rpn]] = l - start_col
                    wr[rpn] += 1
                base = rpn * 2 * kk
                off[base + (j + k - end_row) * k + (l - start_col)] = v
                WV[base + (j + k - end_row) + rr[s] * k] = v
            else:
                start_row = part * rpn + (rpn if rpn < rem else rem)
                end_col = part * cp + (cp + part + 1 if cp < rem else rem + part)
                s = (rpn - 1) * k + l - end_col + k
                if rl[s] < 0:
                    rl[s] = k - 1 - wl[rpn - 1]
                    pl[(rpn - 1) * k + k - 1 - wl[rpn - 1]] = l - end_col + k
                    wl[rpn - 1] += 1
                base = (cp * 2 + 1) * kk
                off[base + (j - start_row) * k + (l - end_col + k)] = v
                WV[base + (j - start_row) + rl[s] * k] = v
    drp = [0] * (n + 1)
    for r in drows:
        drp[r + 1] += 1
    for i in range(n):
        drp[i + 1] += drp[i]
    # second-level reordering
    sre, sp, buf = [0] * n, [0] * n, [0] * n
    nb = eb = 0
    for i in range(P):
        ne = nb + part + (1 if i < rem else 0)
        ee = drp[ne]
        _partitioned_rcm(n, drp, drows, dcols, eb, ee, nb, ne, sre, sp, buf)
        nb, eb = ne, ee
    nr = [sp[r] for r in drows]
    nc = [sp[c] for c in dcols]
    drp2, dci2, dval2 = _regroup_rows_backward(n, nr, nc, dvals)
    first = [0] * (P - 1)
    last_row = 0
    for i in range(P - 1):
        last_row += part + (1 if i < rem else 0)
        first[i] = last_row
        for j in range(last_row - k, last_row):
            if exists[j] and first[i] > sp[j]:
                first[i] = sp[j]
    # assembleBandedMatrix (several partitions)
    ks = [0] * P
    kcol, krow = [0] * n, [0] * n
    for j in range(n):
        for e in range(drp2[j], drp2[j + 1]):
            l = dci2[e]
            cp = _part_of(l, part, rem)
            ks[cp] = max(ks[cp], abs(l - j))
            kcol[l] = max(kcol[l], j - l)
            krow[j] = max(krow[j], l - j)
    bo = [0] * (P + 1)
    for i in range(P):
        bo[i + 1] = bo[i] + (part + (1 if i < rem else 0)) * (2 * ks[i] + 1)
    begin = 0
    for i in range(P):
        end = begin + part + (1 if i < rem else 0)
        for j in range(begin + 1, end):
            kcol[j] = min(max(kcol[j], kcol[j - 1] - 1), end - j - 1)
            krow[j] = min(max(krow[j], krow[j - 1] - 1), end - j - 1)
        begin = end
    f64 = lambda a: np.array(a, dtype=T).astype(np.float64)
    i32 = lambda a: np.array(a, dtype=np.int32)
    return dict(ks=i32(ks), boffsets=np.array(bo, dtype=np.int64), first_rows=i32(first), second_reorder=i32(sre),
                second_perm=i32(sp), widths_left=i32(wl), widths_right=i32(wr), perms_left=i32(pl),
                perms_right=i32(pr), wv=f64(WV), offdiags=f64(off), ks_col=i32(kcol), ks_row=i32(krow),
                drp=i32(drp2), dci=i32(dci2), dval=f64(dval2))
