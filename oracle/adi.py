"""Minimal forward-mode AD with sparse Jacobian blocks -- oracle only.

Restates the semantics of the reference's ADI (pymrst/utils/adi/adi.py:82-223,
init_variables_adi.py:13-24, operators.py:15-72,162-211): a value vector plus
one scipy CSR Jacobian block per primary variable; products use
d(uv) = diag(v) du + diag(u) dv; sparse-matrix products act on the value and
on every block; combine_equations stacks equations row-wise and blocks
column-wise.  Pinned by central finite differences in tests.
"""
import numpy as np
import scipy.sparse as sp


def _rowscale(J, v):
    J = J.tocsr(copy=True)
    J.data *= np.repeat(np.asarray(v, float), np.diff(J.indptr))
    return J


class ADI:
    __array_priority__ = 1000

    def __init__(self, val, jac):
        self.val = np.asarray(val, dtype=float).ravel()
        self.jac = [j.tocsr() for j in jac]

    # -- helpers ---------------------------------------------------------
    def _bcast(self, n):
        """Broadcast a size-1 ADI to n rows."""
        if self.val.size == n:
            return self
        assert self.val.size == 1
        R = sp.csr_array(np.ones((n, 1)))
        return ADI(np.repeat(self.val, n), [R @ j for j in self.jac])

    @staticmethod
    def _pair(a, b):
        if isinstance(a, ADI) and isinstance(b, ADI):
            n = max(a.val.size, b.val.size)
            return a._bcast(n), b._bcast(n)
        return a, b

    def __len__(self):
        return self.val.size

    # -- arithmetic ------------------------------------------------------
    def __neg__(self):
        return ADI(-self.val, [-j for j in self.jac])

    def __add__(self, o):
        if isinstance(o, ADI):
            a, b = ADI._pair(self, o)
            return ADI(a.val + b.val, [x + y for x, y in zip(a.jac, b.jac)])
        o = np.asarray(o, float)
        s = self._bcast(max(self.val.size, o.size))
        return ADI(s.val + o, s.jac)

    __radd__ = __add__

    def __sub__(self, o):
        return self + (-o)

    def __rsub__(self, o):
        return (-self) + o

    def __mul__(self, o):
        if isinstance(o, ADI):
            a, b = ADI._pair(self, o)
            return ADI(a.val * b.val, [_rowscale(x, b.val) + _rowscale(y, a.val)
                                       for x, y in zip(a.jac, b.jac)])
        o = np.asarray(o, float)
        s = self._bcast(max(self.val.size, o.size))
        ov = np.broadcast_to(o, s.val.shape)
        return ADI(s.val * ov, [_rowscale(j, ov) for j in s.jac])

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, ADI):
            a, b = ADI._pair(self, o)
            inv = 1.0 / b.val
            return ADI(a.val * inv, [_rowscale(x, inv) - _rowscale(y, a.val * inv * inv)
                                     for x, y in zip(a.jac, b.jac)])
        return self * (1.0 / np.asarray(o, float))

    def __rtruediv__(self, o):
        o = np.asarray(o, float)
        inv = 1.0 / self.val
        return ADI(o * inv, [_rowscale(j, -o * inv * inv) for j in self.jac])

    def __pow__(self, e):
        e = float(e)
        return ADI(self.val ** e, [_rowscale(j, e * self.val ** (e - 1)) for j in self.jac])

    def __getitem__(self, idx):
        idx = np.asarray(idx)
        if idx.dtype == bool:
            idx = np.flatnonzero(idx)
        S = sp.csr_array((np.ones(idx.size), (np.arange(idx.size), idx)),
                         shape=(idx.size, self.val.size))
        return ADI(self.val[idx], [S @ j for j in self.jac])


def value(x):
    return x.val if isinstance(x, ADI) else np.asarray(x, float)


def ad_exp(x):
    if isinstance(x, ADI):
        e = np.exp(x.val)
        return ADI(e, [_rowscale(j, e) for j in x.jac])
    return np.exp(x)


def ad_matmul(M, x):
    """Sparse matrix times ADI (operators.py:15-72)."""
    if isinstance(x, ADI):
        return ADI(M @ x.val, [(M @ j).tocsr() for j in x.jac])
    return M @ x


def ad_sum(x):
    if isinstance(x, ADI):
        return ad_matmul(sp.csr_array(np.ones((1, x.val.size))), x)
    return np.sum(x)


def init_variables_adi(*vals):
    """init_variables_adi.py:13-24: identity blocks on the diagonal."""
    vals = [np.asarray(v, float).ravel() for v in vals]
    sizes = [v.size for v in vals]
    out = []
    for i, v in enumerate(vals):
        jac = [sp.identity(n, format="csr") if k == i else sp.csr_array((v.size, n))
               for k, n in enumerate(sizes)]
        out.append(ADI(v, jac))
    return out


def combine_equations(*eqs):
    """operators.py:204-211: one value vector, one CSR Jacobian."""
    val = np.concatenate([e.val for e in eqs])
    J = sp.vstack([sp.hstack(e.jac, format="csr") for e in eqs], format="csr")
    return val, J
