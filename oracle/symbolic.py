"""Symbolic analysis for the multifrontal factor -- oracle restatement (test infrastructure only).

Behavioural spec: src/symbolic.jl:15-237.  Indices here are 0-based (the
reference is 1-based; "no parent" is -1 here, 0 there).

  sym_csc        pattern of B + B^T, sorted rows, diagonal dropped   (symbolic.jl:15-33)
  permute_sym    symmetric permutation of such a pattern              (symbolic.jl:40-63)
  etree          Liu's elimination tree with path compression         (symbolic.jl:67-82)
  postorder      DFS postorder, children visited in ascending order   (symbolic.jl:86-116)
  col_structure  union rule: struct(L_j) = A_j(i>j) U children's      (symbolic.jl:122-160)
  supernodes     fundamental supernodes: j joins j-1 iff parent[j-1]==j,
                 j has one child and colcount[j-1] == colcount[j]+1   (symbolic.jl:167-194)
  symbolic_mf    qf = q[post], etree/colstruct/supernodes on A[qf,qf] (symbolic.jl:211-226)
  predicted_fill 2*(n + sum |colstruct[j]|)                           (symbolic.jl:230-237)
"""
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import amd as _amd


def sym_csc(B):
    """(cp, ri) of the pattern of B + B^T, rows ascending, diagonal dropped."""
    P = sp.csc_matrix((np.ones(B.nnz), B.indices.copy(), B.indptr.copy()), shape=B.shape)
    S = (P + P.T).tocsc()
    S.sort_indices()
    n = S.shape[1]
    cp = [0] * (n + 1)
    ri = []
    for j in range(n):
        for p in range(S.indptr[j], S.indptr[j + 1]):
            i = int(S.indices[p])
            if i != j:
                ri.append(i)
        cp[j + 1] = len(ri)
    return cp, ri


def permute_sym(cp, ri, perm):
    """Pattern of B[perm, perm] from the pattern of B (new col j = old col perm[j])."""
    n = len(perm)
    pinv = [0] * n
    for j, o in enumerate(perm):
        pinv[o] = j
    cpN = [0] * (n + 1)
    riN = []
    for j in range(n):
        o = perm[j]
        col = sorted(pinv[ri[p]] for p in range(cp[o], cp[o + 1]))
        riN.extend(col)
        cpN[j + 1] = len(riN)
    return cpN, riN


def etree(cp, ri, n):
    parent = [-1] * n
    ancestor = [-1] * n
    for k in range(n):
        for p in range(cp[k], cp[k + 1]):
            i = ri[p]
            while i != -1 and i < k:
                inext = ancestor[i]
                ancestor[i] = k
                if inext == -1:
                    parent[i] = k
                i = inext
    return parent


def _child_lists(parent):
    n = len(parent)
    head = [-1] * n
    nextc = [-1] * n
    for j in range(n - 1, -1, -1):
        p = parent[j]
        if p != -1:
            nextc[j] = head[p]
            head[p] = j
    return head, nextc


def postorder(parent):
    n = len(parent)
    head, nextc = _child_lists(parent)
    post = []
    for j in range(n):
        if parent[j] != -1:
            continue
        stack = [j]
        while stack:
            p = stack[-1]
            i = head[p]
            if i == -1:
                stack.pop()
                post.append(p)
            else:
                head[p] = nextc[i]
                stack.append(i)
    return post


def col_structure(cp, ri, parent, n):
    head, nextc = _child_lists(parent)
    marker = [-1] * n
    colstruct = [None] * n
    for j in range(n):
        marker[j] = j
        col = []
        for p in range(cp[j], cp[j + 1]):
            i = ri[p]
            if i > j and marker[i] != j:
                marker[i] = j
                col.append(i)
        c = head[j]
        while c != -1:
            for i in colstruct[c]:
                if i > j and marker[i] != j:
                    marker[i] = j
                    col.append(i)
            c = nextc[c]
        col.sort()
        colstruct[j] = col
    return colstruct


def supernodes(parent, colstruct, n):
    colcount = [len(c) for c in colstruct]
    nchild = [0] * n
    for j in range(n):
        if parent[j] != -1:
            nchild[parent[j]] += 1
    super_of = [0] * n
    sstart = []
    ns = -1
    for j in range(n):
        cont = j > 0 and parent[j - 1] == j and nchild[j] == 1 and colcount[j - 1] == colcount[j] + 1
        if not cont:
            ns += 1
            sstart.append(j)
        super_of[j] = ns
    sstart.append(n)
    nsuper = len(sstart) - 1
    childsupers = [[] for _ in range(nsuper)]
    for sk in range(nsuper):
        lastcol = sstart[sk + 1] - 1
        p = parent[lastcol]
        if p != -1:
            childsupers[super_of[p]].append(sk)
    return sstart, super_of, childsupers


@dataclass
class SymbolicMF:
    qf: np.ndarray          # final column order (0-based)
    parent: list            # etree in qf numbering (-1 = root)
    colstruct: list         # colstruct[j] = sorted rows > j of L column j
    sstart: list            # supernode k owns columns sstart[k] .. sstart[k+1]-1
    childsupers: list       # child supernodes of each supernode, ascending

    @property
    def nsuper(self):
        return len(self.sstart) - 1


def symbolic_mf(A, q=None):
    """src/symbolic.jl:211-226.  `q` defaults to the AMD order of A + A^T (0-based)."""
    A = A.tocsc()
    n = A.shape[1]
    qamd = np.asarray(_amd.amd_order_sym(A) if q is None else q, dtype=np.int64)
    B = A[qamd, :][:, qamd].tocsc()
    B.sort_indices()
    cp1, ri1 = sym_csc(B)
    parent1 = etree(cp1, ri1, n)
    post = postorder(parent1)
    qf = qamd[np.asarray(post, dtype=np.int64)]
    cpF, riF = permute_sym(cp1, ri1, post)
    parentF = etree(cpF, riF, n)
    colstruct = col_structure(cpF, riF, parentF, n)
    sstart, _, childsupers = supernodes(parentF, colstruct, n)
    return SymbolicMF(qf, parentF, colstruct, sstart, childsupers)


def predicted_fill(S):
    n = len(S.qf)
    return 2 * (n + sum(len(c) for c in S.colstruct))


def dense_symbolic_fill(B):
    """nnz(L)+nnz(U) of a dense no-cancellation elimination of pattern(B + B^T) in natural order.

    Brute force (O(n^3) booleans) used to pin `col_structure` on small inputs.
    """
    n = B.shape[0]
    M = (B.toarray() != 0)
    M = M | M.T
    np.fill_diagonal(M, True)
    nl = 0
    for k in range(n):
        rows = np.nonzero(M[k + 1:, k])[0] + k + 1
        nl += len(rows)
        if len(rows):
            M[np.ix_(rows, rows)] = True
    return 2 * (n + nl)
