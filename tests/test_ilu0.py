"""The ILU(0) oracle (MRST's nofill ILU of gmres_ilu_solver_ad.py) pinned by
what fixes ILU(0) uniquely: L unit lower, U upper, both on A's pattern, and
(L U)_ij = A_ij on every position of the pattern; a full pattern gives the
exact LU; lu_solve inverts L U.  A planted mistake (the update skipped) is
caught.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle.ilu as oilu
from mutation import fails, planted


def _block_tpfa(n_side=4, B=2, seed=0):
    """A 2-D TPFA-patterned matrix with dense B x B blocks (stored zeros
    kept), diagonally dominant, in cell-major interleaved order."""
    rng = np.random.default_rng(seed)
    nc = n_side * n_side
    rows, cols = [], []
    for c in range(nc):
        i, j = c % n_side, c // n_side
        nb = [c] + [c + d for d, ok in ((1, i + 1 < n_side), (-1, i > 0), (n_side, j + 1 < n_side),
                                          (-n_side, j > 0)) if ok]
        for m in nb:
            for a in range(B):
                for b in range(B):
                    rows.append(c * B + a)
                    cols.append(m * B + b)
    vals = rng.normal(size=len(rows))
    A = sp.coo_array((vals, (rows, cols)), shape=(nc * B, nc * B)).tocsr()
    A = A + sp.diags_array(np.full(nc * B, 8.0))
    return sp.csr_array(A)


def check_ilu0_is_the_unique_pattern_factorisation():
    A = _block_tpfa()
    L, U = oilu.ilu0(A)
    mask = sp.csr_array((np.ones(A.nnz), A.indices, A.indptr), shape=A.shape).toarray() != 0
    LU = (L @ U).toarray()
    Ad = A.toarray()
    assert np.abs(LU - Ad)[mask].max() <= 1e-12 * np.abs(Ad).max()
    assert np.abs(np.triu(L.toarray(), 1)).max() == 0 and np.allclose(np.diag(L.toarray()), 1.0)
    assert np.abs(np.tril(U.toarray(), -1)).max() == 0
    assert not (np.abs(L.toarray()) > 0)[~mask].any() and not (np.abs(U.toarray()) > 0)[~mask].any()
    # a full pattern: ILU(0) is the exact LU
    D = sp.csr_array(np.random.default_rng(1).normal(size=(6, 6)) + 6 * np.eye(6))
    L, U = oilu.ilu0(D)
    assert np.abs((L @ U).toarray() - D.toarray()).max() <= 1e-13 * 6
    r = np.arange(6.0)
    assert np.abs(oilu.lu_solve(L, U, r) - np.linalg.solve(D.toarray(), r)).max() <= 1e-12


def test_ilu0_known_answers():
    check_ilu0_is_the_unique_pattern_factorisation()


def test_ilu0_check_sees_a_planted_mistake():
    with planted(oilu, "ilu0", ("val[pos] -= val[p] * val[q]", "val[pos] -= 0.0 * val[q]")):
        assert fails(check_ilu0_is_the_unique_pattern_factorisation)
