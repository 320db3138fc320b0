"""The twisted block-tridiagonal factorization that the line local stage of
CPR uses (k_line_factor_tw / k_line_gs_tw in pymrst_amd/csrc/linsys.hip),
restated in NumPy and checked against a dense solve for every line length
from 1 to 8 and block sizes 2 and 3 (CPU only).

The top half [0, m) of a line is eliminated downwards,
D'_k = D_k - L_k D'_{k-1}^{-1} U_{k-1}, the bottom half (m, L) upwards,
D''_k = D_k - U_k D''_{k+1}^{-1} L_{k+1}, with m = L // 2; the middle pivot
takes both couplings; the right-hand side is walked the same way and the
solution is recovered outwards from the middle.  A wrong middle index or a
swapped coupling shows up as a wrong solution here."""
import numpy as np
import pytest


def twisted_solve(D, Lo, Up, r):
    """D[k] diagonal blocks, Lo[k] the coupling of cell k to k-1, Up[k] of
    cell k to k+1, r[k] right-hand sides; returns x with A x = r."""
    L = len(D)
    m = L // 2
    inv = [None] * L
    y = [None] * L
    for k in range(m):                       # top half, downwards
        Dk = D[k] - (Lo[k] @ inv[k - 1] @ Up[k - 1] if k > 0 else 0.0)
        inv[k] = np.linalg.inv(Dk)
        y[k] = inv[k] @ (r[k] - (Lo[k] @ y[k - 1] if k > 0 else 0.0))
    for k in range(L - 1, m, -1):            # bottom half, upwards
        Dk = D[k] - (Up[k] @ inv[k + 1] @ Lo[k + 1] if k < L - 1 else 0.0)
        inv[k] = np.linalg.inv(Dk)
        y[k] = inv[k] @ (r[k] - (Up[k] @ y[k + 1] if k < L - 1 else 0.0))
    Dm = D[m].copy()
    rm = r[m].copy()
    if m > 0:
        Dm = Dm - Lo[m] @ inv[m - 1] @ Up[m - 1]
        rm = rm - Lo[m] @ y[m - 1]
    if m < L - 1:
        Dm = Dm - Up[m] @ inv[m + 1] @ Lo[m + 1]
        rm = rm - Up[m] @ y[m + 1]
    x = [None] * L
    x[m] = np.linalg.solve(Dm, rm)
    for k in range(m - 1, -1, -1):           # back out, upwards on the top half
        x[k] = y[k] - inv[k] @ (Up[k] @ x[k + 1])
    for k in range(m + 1, L):                # and downwards on the bottom half
        x[k] = y[k] - inv[k] @ (Lo[k] @ x[k - 1])
    return np.concatenate(x)


def _random_line(L, B, rng):
    D = [np.eye(B) * (4.0 + B) + rng.normal(0, 0.5, (B, B)) for _ in range(L)]
    Lo = [rng.normal(0, 1.0, (B, B)) for _ in range(L)]
    Up = [rng.normal(0, 1.0, (B, B)) for _ in range(L)]
    A = np.zeros((L * B, L * B))
    for k in range(L):
        A[k * B:(k + 1) * B, k * B:(k + 1) * B] = D[k]
        if k > 0:
            A[k * B:(k + 1) * B, (k - 1) * B:k * B] = Lo[k]
        if k < L - 1:
            A[k * B:(k + 1) * B, (k + 1) * B:(k + 2) * B] = Up[k]
    return D, Lo, Up, A


@pytest.mark.parametrize("B", [2, 3])
@pytest.mark.parametrize("L", list(range(1, 9)))
def test_twisted_solve_equals_dense_solve(L, B):
    rng = np.random.default_rng(100 * B + L)
    D, Lo, Up, A = _random_line(L, B, rng)
    r = rng.normal(size=L * B)
    x = twisted_solve(D, Lo, Up, list(r.reshape(L, B)))
    assert np.abs(x - np.linalg.solve(A, r)).max() <= 1e-13 * np.abs(x).max()


def test_a_wrong_middle_fails():
    """The same walk with the middle cell taken one lower is not a solve."""
    rng = np.random.default_rng(7)
    D, Lo, Up, A = _random_line(6, 2, rng)
    r = rng.normal(size=12)
    Dw = D[::-1]
    # reversing the line and its couplings moves the meeting point from 3 to 2
    # in the original order; the reversal itself is exact:
    x_rev = twisted_solve(Dw, Up[::-1], Lo[::-1], list(r.reshape(6, 2)[::-1]))
    assert np.abs(x_rev.reshape(6, 2)[::-1].ravel() - np.linalg.solve(A, r)).max() <= 1e-12
    # but swapping the two couplings (a transposed operand) is caught
    x_bad = twisted_solve(D, Up, Lo, list(r.reshape(6, 2)))
    assert np.abs(x_bad - np.linalg.solve(A, r)).max() > 1e-3


# ---------------------------------------------------------------------------
# Partitioned line solves (k_line_factor_seg / k_line_reduce / k_line_gs_seg):
# a line of L >= 32 cells is cut at ns - 1 separators into ns segments, each
# solved by the twisted walk with the separators held at zero; the spikes
# give every segment cell's response to its separators, the separators solve
# their block-tridiagonal Schur system by block Thomas, and each cell is
# corrected by x = g - W_L x_left - W_R x_right.

def line_nseg(L):
    n = L // 16
    return 8 if n >= 8 else 4 if n >= 4 else 2 if n >= 2 else 1


def line_sep(L, ns, k):
    return (k + 1) * (L + 1) // ns - 1


def line_segment(L, ns, k):
    a = 0 if k == 0 else line_sep(L, ns, k - 1) + 1
    b = L - 1 if k == ns - 1 else line_sep(L, ns, k) - 1
    return a, b


def _segment_factors(D, Lo, Up, a, b):
    """Twisted factors of cells [a, b] with the walk couplings as the kernel
    stores them: inv, Lo' (towards the previous cell of the half's walk) and
    P = inv Up' off the middle, the raw couplings at the middle."""
    m = a + (b - a + 1) // 2
    Z = np.zeros_like(D[a])
    inv, lo_, up_ = {}, {}, {}
    for k in range(a, m):
        Dk = D[k] - (Lo[k] @ inv[k - 1] @ Up[k - 1] if k > a else 0.0)
        inv[k] = np.linalg.inv(Dk)
        lo_[k] = Lo[k] if k > a else Z
        up_[k] = inv[k] @ Up[k]
    for k in range(b, m, -1):
        Dk = D[k] - (Up[k] @ inv[k + 1] @ Lo[k + 1] if k < b else 0.0)
        inv[k] = np.linalg.inv(Dk)
        lo_[k] = Up[k] if k < b else Z
        up_[k] = inv[k] @ Lo[k]
    Dm = D[m].copy()
    if m > a:
        Dm = Dm - Lo[m] @ inv[m - 1] @ Up[m - 1]
    if m < b:
        Dm = Dm - Up[m] @ inv[m + 1] @ Lo[m + 1]
    inv[m] = np.linalg.inv(Dm)
    lo_[m] = Lo[m] if m > a else Z
    up_[m] = Up[m] if m < b else Z
    return m, inv, lo_, up_


def _segment_walk(a, b, m, inv, lo_, up_, rhs):
    """The walk of k_line_gs_seg (and of the spike solves) on [a, b]: forward
    to the middle from both ends, then g = y - P g_next back out; rhs maps a
    cell to its right-hand side (vectors or B x B blocks)."""
    y = {}
    prev = np.zeros_like(rhs[a])
    for k in range(a, m):
        prev = y[k] = inv[k] @ (rhs[k] - lo_[k] @ prev)
    ytop = prev
    prev = np.zeros_like(rhs[a])
    for k in range(b, m, -1):
        prev = y[k] = inv[k] @ (rhs[k] - lo_[k] @ prev)
    ybot = prev
    g = {m: inv[m] @ (rhs[m] - lo_[m] @ ytop - up_[m] @ ybot)}
    for k in range(m - 1, a - 1, -1):
        g[k] = y[k] - up_[k] @ g[k + 1]
    for k in range(m + 1, b + 1):
        g[k] = y[k] - up_[k] @ g[k - 1]
    return g


def partitioned_solve(D, Lo, Up, r):
    L = len(D)
    ns = line_nseg(L)
    if ns == 1:
        return twisted_solve(D, Lo, Up, r)
    segs = [line_segment(L, ns, k) for k in range(ns)]
    fac = [_segment_factors(D, Lo, Up, a, b) for a, b in segs]
    Z = np.zeros_like(D[0])
    WL, WR, g = [], [], []
    for k, ((a, b), (m, inv, lo_, up_)) in enumerate(zip(segs, fac)):
        zero = {c: Z for c in range(a, b + 1)}
        WL.append(_segment_walk(a, b, m, inv, lo_, up_, {**zero, a: Lo[a]}) if k > 0 else zero)
        WR.append(_segment_walk(a, b, m, inv, lo_, up_, {**zero, b: Up[b]}) if k < ns - 1 else zero)
        g.append(_segment_walk(a, b, m, inv, lo_, up_, {c: r[c] for c in range(a, b + 1)}))
    # separators: block Thomas on the Schur system
    seps = [line_sep(L, ns, k) for k in range(ns - 1)]
    Dinv, Tsub, Ustore, zk = [], [], [], []
    Uprev = Z
    zp = np.zeros_like(r[0])
    for k, s in enumerate(seps):
        (_, bk), (ak1, _) = segs[k], segs[k + 1]
        Tkk = D[s] - Lo[s] @ WR[k][bk] - Up[s] @ WL[k + 1][ak1]
        Ts = -(Lo[s] @ WL[k][bk])
        Tp = -(Up[s] @ WR[k + 1][ak1])
        Di = np.linalg.inv(Tkk - Ts @ Uprev)
        Uprev = Di @ Tp
        rhs = r[s] - Lo[s] @ g[k][bk] - Up[s] @ g[k + 1][ak1] - Ts @ zp
        zp = Di @ rhs
        zk.append(zp)
        Ustore.append(Uprev)
    xs = [None] * len(seps)
    xn = np.zeros_like(r[0])
    for k in range(len(seps) - 1, -1, -1):
        xn = xs[k] = zk[k] - Ustore[k] @ xn
    x = [None] * L
    for k, (a, b) in enumerate(segs):
        xl = xs[k - 1] if k > 0 else np.zeros_like(r[0])
        xr = xs[k] if k < ns - 1 else np.zeros_like(r[0])
        for c in range(a, b + 1):
            x[c] = g[k][c] - WL[k][c] @ xl - WR[k][c] @ xr
    for k, s in enumerate(seps):
        x[s] = xs[k]
    return np.concatenate(x)


def test_segment_geometry_covers_every_cell_once():
    for L in range(1, 400):
        ns = line_nseg(L)
        seen = np.zeros(L, dtype=int)
        for k in range(ns):
            a, b = line_segment(L, ns, k)
            assert b - a + 1 >= 15 or ns == 1, (L, ns, k)
            seen[a:b + 1] += 1
        for k in range(ns - 1):
            seen[line_sep(L, ns, k)] += 1
        assert (seen == 1).all(), L
    assert [line_nseg(L) for L in (31, 32, 63, 64, 85, 127, 128, 300)] == [1, 2, 2, 4, 4, 4, 8, 8]


@pytest.mark.parametrize("B", [2, 3])
@pytest.mark.parametrize("L", [1, 2, 31, 32, 33, 47, 63, 64, 85, 100, 127, 128, 161, 200])
def test_partitioned_solve_equals_dense_solve(L, B):
    rng = np.random.default_rng(1000 * B + L)
    D, Lo, Up, A = _random_line(L, B, rng)
    r = rng.normal(size=L * B)
    x = partitioned_solve(D, Lo, Up, list(r.reshape(L, B)))
    assert np.abs(x - np.linalg.solve(A, r)).max() <= 1e-12 * np.abs(x).max()


def test_a_dropped_spike_term_fails():
    """Leaving out the right spike's correction is not a solve."""
    rng = np.random.default_rng(11)
    L, B = 85, 2
    D, Lo, Up, A = _random_line(L, B, rng)
    r = rng.normal(size=L * B)
    Upz = [u.copy() for u in Up]
    s = line_sep(L, 4, 1)
    Upz[s - 1] = np.zeros((B, B))     # the coupling into separator 1 forgotten
    x = partitioned_solve(D, Lo, Upz, list(r.reshape(L, B)))
    assert np.abs(x - np.linalg.solve(A, r)).max() > 1e-3
