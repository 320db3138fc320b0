"""ILU(0) oracle (test infrastructure only).

Restates MRST's ILU(0) of the reference, pymrst/modules/ad_core/solvers/
gmres_ilu_solver_ad.py: ilu :119-138 (L unit lower, U upper, A's pattern),
_ilu0 :141-170 (in-place IKJ elimination over sorted CSR rows, zero fill),
_forward / _backward :178-209 (the triangular solves).  ILU(0) is unique for
a pattern: L U equals A on every position of the pattern (:13-15).
"""
import numpy as np
import scipy.sparse as sp


def ilu0(A):
    """(L, U) of A's ILU(0); every stored entry of A, zeros included, is in
    the pattern."""
    A = sp.csr_array(A)
    A.sort_indices()
    n = A.shape[0]
    ptr, idx = A.indptr, A.indices
    val = A.data.astype(float).copy()
    diag = np.full(n, -1)
    for i in range(n):
        hit = np.flatnonzero(idx[ptr[i]:ptr[i + 1]] == i)
        if hit.size == 0:
            raise ZeroDivisionError(f"ILU(0): row {i} has no diagonal entry")
        diag[i] = ptr[i] + hit[0]
    where = {}
    for i in range(n):
        where = {int(idx[p]): p for p in range(ptr[i], ptr[i + 1])}
        for p in range(ptr[i], diag[i]):
            k = idx[p]
            if val[diag[k]] == 0.0:
                raise ZeroDivisionError(f"ILU(0): zero pivot in row {k}")
            val[p] /= val[diag[k]]
            for q in range(diag[k] + 1, ptr[k + 1]):
                pos = where.get(int(idx[q]))
                if pos is not None:
                    val[pos] -= val[p] * val[q]
    F = sp.csr_array((val, idx.copy(), ptr.copy()), shape=A.shape)
    L = sp.tril(F, k=-1, format="csr") + sp.eye_array(n, format="csr")
    U = sp.triu(F, format="csr")
    return sp.csr_array(L), sp.csr_array(U)


def lu_solve(L, U, r):
    """x with L U x = r (forward then backward substitution)."""
    L, U = sp.csr_array(L), sp.csr_array(U)
    y = np.asarray(r, float).ravel().copy()
    n = y.size
    for i in range(n):
        row = slice(L.indptr[i], L.indptr[i + 1])
        j, v = L.indices[row], L.data[row]
        m = j < i
        y[i] -= v[m] @ y[j[m]]
    x = y.copy()
    for i in range(n - 1, -1, -1):
        row = slice(U.indptr[i], U.indptr[i + 1])
        j, v = U.indices[row], U.data[row]
        m = j > i
        x[i] = (x[i] - v[m] @ x[j[m]]) / v[j == i][0]
    return x
