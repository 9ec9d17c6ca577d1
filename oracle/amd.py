"""Approximate minimum degree ordering -- oracle restatement (test infrastructure only).

Behavioural spec: src/amd.jl:84-188 (`_sym_pattern`, `_ata_pattern`,
`amd_order_sym`, `colamd_order`) and :191-539 (`_amd_core`), which follow
CSparse's `cs_amd` (Davis, "Direct Methods for Sparse Linear Systems", SIAM
2006, section 7.1; Amestoy, Davis & Duff, SIAM J. Matrix Anal. Appl. 17(4),
1996).  The ordering is only reproducible step-for-step, so this restatement
keeps the published algorithm's phases in the same order: quotient-graph
elimination with element absorption, approximate external degrees from the
|Le \\ Lk| set differences, hash-based supervariable detection, aggressive
absorption, mass elimination, and a final postorder of the assembly tree.

Indices are 0-based; FLIP(i) = -i-2 encodes "absorbed into i".
Pinned by the published fill counts in results/SUMMARY.md:32-39 (tests).
"""
import math

import numpy as np
import scipy.sparse as sp


def _flip(i):
    return -i - 2


def _pattern_offdiag(S):
    """0-based CSC (cp, ci) of S's stored pattern with the diagonal removed (src/amd.jl:52-75)."""
    S = S.tocsc()
    S.sort_indices()
    n = S.shape[1]
    cp = [0] * (n + 1)
    ci = []
    for j in range(n):
        for p in range(S.indptr[j], S.indptr[j + 1]):
            i = int(S.indices[p])
            if i != j:
                ci.append(i)
        cp[j + 1] = len(ci)
    return cp, ci


def _pattern_matrix(A):
    A = A.tocsc()
    return sp.csc_matrix((np.ones(A.nnz), A.indices.copy(), A.indptr.copy()), shape=A.shape)


def sym_pattern(A):
    """Pattern of A + A^T, diagonal dropped (src/amd.jl:84-87)."""
    P = _pattern_matrix(A)
    return _pattern_offdiag((P + P.T).tocsc())


def ata_pattern(A, dense_param=10.0):
    """Pattern of A^T A with dense rows of A removed, diagonal dropped (src/amd.jl:91-104)."""
    n = A.shape[1]
    P = _pattern_matrix(A)
    AT = P.T.tocsc()                       # column j of AT = row j of A
    rowcnt = np.diff(AT.indptr)
    thr = int(math.floor(min(n - 2, max(16.0, dense_param * math.sqrt(n)))))
    keep = rowcnt <= thr
    AT2 = AT[:, np.nonzero(keep)[0]]
    C = (AT2 @ AT2.T).tocsc()
    return _pattern_offdiag(C)


def _wclear(mark, lemax, w, n):
    if mark < 2 or mark + lemax < 0:
        for k in range(n):
            if w[k] != 0:
                w[k] = 1
        mark = 2
    return mark


def _tdfs(j, k, head, nxt, post, stack):
    top = 0
    stack[0] = j
    while top >= 0:
        p = stack[top]
        i = head[p]
        if i == -1:
            top -= 1
            post[k] = p
            k += 1
        else:
            head[p] = nxt[i]
            top += 1
            stack[top] = i
    return k


def amd_core(n, cp, ci, dense=10.0):
    """Quotient-graph AMD on a symmetric 0-based pattern; returns a 0-based permutation."""
    cnz = cp[n]
    nzmax = cnz + cnz // 5 + 2 * n
    Ci = list(ci) + [0] * (max(nzmax, 1) - cnz)
    nzmax = len(Ci)
    dthr = max(16.0, dense * math.sqrt(n))
    densethr = min(n - 2, int(math.floor(dthr)))

    ln_ = [cp[k + 1] - cp[k] for k in range(n)] + [0]       # len
    Cp = list(cp)
    nv = [1] * (n + 1)
    nxt = [-1] * (n + 1)
    head = [-1] * (n + 1)
    elen = [0] * (n + 1)
    degree = list(ln_)
    w = [1] * (n + 1)
    hhead = [-1] * (n + 1)
    last = [-1] * (n + 1)
    P = [0] * (n + 1)

    mark = _wclear(0, 0, w, n)
    elen[n] = -2
    Cp[n] = -1
    w[n] = 0
    nel = 0
    for i in range(n):
        d = degree[i]
        if d == 0:
            elen[i] = -2
            nel += 1
            Cp[i] = -1
            w[i] = 0
        elif d > densethr:
            nv[i] = 0
            elen[i] = -1
            nel += 1
            Cp[i] = _flip(n)
            nv[n] += 1
        else:
            if head[d] != -1:
                last[head[d]] = i
            nxt[i] = head[d]
            head[d] = i

    mindeg = 0
    lemax = 0
    while nel < n:
        # pivot of minimum approximate degree
        k = -1
        while mindeg < n:
            k = head[mindeg]
            if k != -1:
                break
            mindeg += 1
        if nxt[k] != -1:
            last[nxt[k]] = -1
        head[mindeg] = nxt[k]
        elenk = elen[k]
        nvk = nv[k]
        nel += nvk

        # compact the element/variable lists when the free tail is too short
        if elenk > 0 and cnz + mindeg >= nzmax:
            for j in range(n):
                p = Cp[j]
                if p >= 0:
                    Cp[j] = Ci[p]
                    Ci[p] = _flip(j)
            q = 0
            p = 0
            while p < cnz:
                j = _flip(Ci[p])
                p += 1
                if j >= 0:
                    Ci[q] = Cp[j]
                    Cp[j] = q
                    q += 1
                    for _ in range(ln_[j] - 1):
                        Ci[q] = Ci[p]
                        q += 1
                        p += 1
            cnz = q

        # new element Lk = union of k's elements and variables
        dk = 0
        nv[k] = -nvk
        p = Cp[k]
        pk1 = p if elenk == 0 else cnz
        pk2 = pk1
        for k1 in range(1, elenk + 2):
            if k1 > elenk:
                e = k
                pj = p
                ln = ln_[k] - elenk
            else:
                e = Ci[p]
                p += 1
                pj = Cp[e]
                ln = ln_[e]
            for _ in range(ln):
                i = Ci[pj]
                pj += 1
                nvi = nv[i]
                if nvi <= 0:
                    continue
                dk += nvi
                nv[i] = -nvi
                Ci[pk2] = i
                pk2 += 1
                if nxt[i] != -1:
                    last[nxt[i]] = last[i]
                if last[i] != -1:
                    nxt[last[i]] = nxt[i]
                else:
                    head[degree[i]] = nxt[i]
            if e != k:
                Cp[e] = _flip(k)
                w[e] = 0
        if elenk != 0:
            cnz = pk2
        degree[k] = dk
        Cp[k] = pk1
        ln_[k] = pk2 - pk1
        elen[k] = -2

        # |Le \ Lk| for every element e adjacent to a variable of Lk
        mark = _wclear(mark, lemax, w, n)
        for pk in range(pk1, pk2):
            i = Ci[pk]
            eln = elen[i]
            if eln <= 0:
                continue
            nvi = -nv[i]
            wnvi = mark - nvi
            for p in range(Cp[i], Cp[i] + eln):
                e = Ci[p]
                if w[e] >= mark:
                    w[e] -= nvi
                elif w[e] != 0:
                    w[e] = degree[e] + wnvi

        # approximate degrees, element absorption, hash of each variable
        for pk in range(pk1, pk2):
            i = Ci[pk]
            p1 = Cp[i]
            p2 = p1 + elen[i] - 1
            pn = p1
            h = 0
            d = 0
            for p in range(p1, p2 + 1):
                e = Ci[p]
                if w[e] != 0:
                    dext = w[e] - mark
                    if dext > 0:
                        d += dext
                        Ci[pn] = e
                        pn += 1
                        h += e
                    else:
                        Cp[e] = _flip(k)
                        w[e] = 0
            elen[i] = pn - p1 + 1
            p3 = pn
            p4 = p1 + ln_[i]
            for p in range(p2 + 1, p4):
                j = Ci[p]
                nvj = nv[j]
                if nvj <= 0:
                    continue
                d += nvj
                Ci[pn] = j
                pn += 1
                h += j
            if d == 0:
                Cp[i] = _flip(k)
                nvi = -nv[i]
                dk -= nvi
                nvk += nvi
                nel += nvi
                nv[i] = 0
                elen[i] = -1
            else:
                degree[i] = min(degree[i], d)
                Ci[pn] = Ci[p3]
                Ci[p3] = Ci[p1]
                Ci[p1] = k
                ln_[i] = pn - p1 + 1
                h = h % n
                nxt[i] = hhead[h]
                hhead[h] = i
                last[i] = h
        degree[k] = dk
        lemax = max(lemax, dk)
        mark = _wclear(mark + lemax, lemax, w, n)

        # supervariables: variables in the same hash bucket with identical lists
        for pk in range(pk1, pk2):
            i = Ci[pk]
            if nv[i] >= 0:
                continue
            h = last[i]
            i = hhead[h]
            hhead[h] = -1
            while i != -1 and nxt[i] != -1:
                ln = ln_[i]
                eln = elen[i]
                for p in range(Cp[i] + 1, Cp[i] + ln):
                    w[Ci[p]] = mark
                jlast = i
                j = nxt[i]
                while j != -1:
                    ok = ln_[j] == ln and elen[j] == eln
                    p = Cp[j] + 1
                    while ok and p <= Cp[j] + ln - 1:
                        if w[Ci[p]] != mark:
                            ok = False
                        p += 1
                    if ok:
                        Cp[j] = _flip(i)
                        nv[i] += nv[j]
                        nv[j] = 0
                        elen[j] = -1
                        j = nxt[j]
                        nxt[jlast] = j
                    else:
                        jlast = j
                        j = nxt[j]
                i = nxt[i]
                mark += 1

        # finalise Lk and put its variables back into the degree lists
        p = pk1
        for pk in range(pk1, pk2):
            i = Ci[pk]
            nvi = -nv[i]
            if nvi <= 0:
                continue
            nv[i] = nvi
            d = degree[i] + dk - nvi
            d = min(d, n - nel - nvi)
            if head[d] != -1:
                last[head[d]] = i
            nxt[i] = head[d]
            last[i] = -1
            head[d] = i
            mindeg = min(mindeg, d)
            degree[i] = d
            Ci[p] = i
            p += 1
        nv[k] = nvk
        ln_[k] = p - pk1
        if ln_[k] == 0:
            Cp[k] = -1
            w[k] = 0
        if elenk != 0:
            cnz = p

    # postorder the assembly tree
    for i in range(n):
        Cp[i] = _flip(Cp[i])
    for j in range(n + 1):
        head[j] = -1
    for j in range(n, -1, -1):
        if nv[j] > 0:
            continue
        nxt[j] = head[Cp[j]]
        head[Cp[j]] = j
    for e in range(n, -1, -1):
        if nv[e] <= 0:
            continue
        if Cp[e] != -1:
            nxt[e] = head[Cp[e]]
            head[Cp[e]] = e
    k = 0
    for i in range(n + 1):
        if Cp[i] == -1:
            k = _tdfs(i, k, head, nxt, P, w)
    return np.array(P[:n], dtype=np.int64)


def amd_order_sym(A, dense=10.0):
    """0-based AMD permutation of the pattern of A + A^T (src/amd.jl:141-148)."""
    n = A.shape[1]
    if n == 0:
        return np.zeros(0, dtype=np.int64)
    if n == 1:
        return np.zeros(1, dtype=np.int64)
    cp, ci = sym_pattern(A)
    return amd_core(n, cp, ci, dense)


def colamd_order(A, dense=10.0):
    """0-based AMD permutation of the A^T A pattern, dense rows dropped (src/amd.jl:183-188)."""
    n = A.shape[1]
    if n <= 1:
        return np.arange(n, dtype=np.int64)
    cp, ci = ata_pattern(A, dense)
    return amd_core(n, cp, ci, dense)
