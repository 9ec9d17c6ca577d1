"""Row scaling restated from src/scaling.jl:92-126 (oracle; test infrastructure only).

Rs[i] = 1 / sum_j |a_ij| (SUM) or 1 / max_j |a_ij| (MAX); a row with no
nonzero magnitude gets 1; NONE gives all ones.  The sums run over the stored
entries in CSC order (src/scaling.jl:101-103), which numpy's add.at reproduces
exactly (sequential accumulation in storage order).
"""
import numpy as np

SCALE_NONE, SCALE_SUM, SCALE_MAX = 0, 1, 2


def row_scaling(A, kind: int) -> np.ndarray:
    """src/scaling.jl:92-115."""
    n = A.shape[0]
    Rs = np.ones(n, dtype=np.float64)
    if kind == SCALE_NONE:
        return Rs
    acc = np.zeros(n, dtype=np.float64)
    a = np.abs(A.data)
    if kind == SCALE_SUM:
        for p in range(A.nnz):          # sequential order, as the reference loop
            acc[A.indices[p]] += a[p]
    else:
        np.maximum.at(acc, A.indices, a)
    nz = acc > 0
    Rs[nz] = 1.0 / acc[nz]
    return Rs


def apply_row_scaling(A, Rs):
    """src/scaling.jl:118-126: copy of A with row i multiplied by Rs[i] (same pattern)."""
    B = A.copy()
    B.data = B.data * Rs[B.indices]
    return B
