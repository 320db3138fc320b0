"""The linear-solver seam's host side (CPU): the reference's variable-blocked
Jacobian split into the device's cell blocks and per-perforation well
blocks, checked by reassembling them into a dense matrix in device order
(oil-water with a cross-flowing well, whose wellbore mixture couples
perforated cells; black-oil with the same), and the requirements check."""
from types import SimpleNamespace

import numpy as np
import pytest

import oracle.grid as og
from cases import build_oracle, small_gravity_inputs


def _dense_from_pieces(nc, B, pieces):
    rowptr, col, val, perf_ptr, perf_cells, Bp, Cp, Dp, _ = pieces
    M = B + 1
    nw = len(perf_ptr) - 1
    n = nc * B + nw * M
    A = np.zeros((n, n))
    for r in range(nc):
        for k in range(rowptr[r], rowptr[r + 1]):
            A[r * B:(r + 1) * B, col[k] * B:(col[k] + 1) * B] += val[k]
    for w in range(nw):
        o = nc * B + w * M
        A[o:o + M, o:o + M] += Dp[w]
        for j in range(perf_ptr[w], perf_ptr[w + 1]):
            c = perf_cells[j]
            A[c * B:(c + 1) * B, o:o + M] += Bp[j]
            A[o:o + M, c * B:(c + 1) * B] += Cp[j]
    return A


def _check(J, nc, B):
    from pymrst_amd.linear import split_reference_jacobian
    pieces = split_reference_jacobian(J, nc, B)
    dev = pieces[-1]
    Jd = np.zeros(J.shape)
    a = J.tocoo()
    np.add.at(Jd, (dev[a.row], dev[a.col]), a.data)
    A = _dense_from_pieces(nc, B, pieces)
    assert np.array_equal(A, Jd)
    rowptr, col = pieces[0], pieces[1]
    for r in range(nc):      # every diagonal block, columns ascending
        cols = col[rowptr[r]:rowptr[r + 1]]
        assert r in cols and np.all(np.diff(cols) > 0)
    return pieces


def test_oil_water_jacobian_split_reassembles_exactly():
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    st = so.copy()
    rng = np.random.default_rng(3)
    st.pressure = st.pressure + rng.normal(0, 2e5, st.pressure.size)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, st.sW.size), 0.0, 1.0)
    # the third well's perforations straddle its bhp: some inject, some produce
    w = mo.wells[2]
    st.pressure[w.cells] = st.bhp[2] + 3e5 * np.where(np.arange(w.cells.size) % 2 == 0, 1.0, -1.0)
    cdp = mo.compute_cdp_per_well(so)
    _, J, _ = mo.residual_and_jacobian(so, st, 10 * og.DAY, cdp)
    pieces = _check(J, mo.nc, 2)
    # the cross-flowing well's mixture couples its perforated cells beyond
    # the TPFA neighbours (its bottom cell to its top cell): a cell block now
    rowptr, col = pieces[0], pieces[1]
    top, bottom = int(w.cells[0]), int(w.cells[-1])
    assert top in col[rowptr[bottom]:rowptr[bottom + 1]]


def test_black_oil_jacobian_split_reassembles_exactly():
    from oracle.bo import BOFluid, BOModel, synthetic_deck
    dims = (4, 3, 3)
    nx, ny, nz = dims
    g = og.cart_grid(list(dims), [30.0 * nx, 30.0 * ny, 5.0 * nz])
    rng = np.random.default_rng(0)
    perm = 100 * og.MILLI * og.DARCY * np.exp(0.5 * rng.standard_normal(g.num_cells))
    rock = og.make_rock(g, perm, 0.2 + 0.05 * rng.random(g.num_cells))
    N, T = og.internal_connections(g, og.compute_trans(g, rock))[:2]
    nc = g.num_cells
    ws = og.add_well(None, g, rock, [nx * ny * k for k in range(nz)], "bhp", 200 * og.BARSA, compi=[0.0, 0.0, 1.0])
    ws = og.add_well(ws, g, rock, [nc - 1 - nx * ny, nc - 1], "rate", -3e-4, compi=[0.0, 1.0, 0.0])
    m = BOModel(g, rock, BOFluid(synthetic_deck()), ws, N, T, gravity=og.NORM_GRAVITY)
    st = m.initial_state(200 * og.BARSA, 0.25, 0.0)
    st.pressure = st.pressure + rng.normal(0, 5e5, nc)
    gas = rng.random(nc) < 0.5
    sg = np.where(gas, 0.15, 0.0)
    st.s = np.column_stack([np.full(nc, 0.3), 0.7 - sg, sg])
    rsat = m.fluid.rs_sat(st.pressure)
    st.rs = np.where(gas, rsat, 0.8 * rsat)
    st.status = np.where(gas, 3, 1)
    w = m.wells[0]
    st.pressure[w.cells] = st.bhp[0] + 8e5 * np.where(np.arange(w.cells.size) % 2 == 0, 1.0, -1.0)
    _, J = m.residual_and_jacobian(st, st, 10 * og.DAY)[:2]
    _check(J, nc, 3)


def test_requirements_and_the_ad_oo_problem_plumbing():
    """The reference's requirement list (linear_solvers.py:192-224), the
    AD-OO problem's right-hand side (minus the stored residual) and its split
    of the increment per primary variable (storeIncrements), from the
    declared sizes and from AD equations' Jacobian blocks."""
    import scipy.sparse as sp
    from pymrst_amd.linear import LinearizedProblemAD, problem_system, split_increment, unmet_requirements
    assert unmet_requirements(None) == ["no model was supplied"]
    assert unmet_requirements(SimpleNamespace(G=None, n_phases=2)) == ["model has no .fluid"]
    fl = SimpleNamespace(disgas=False, b=None)
    assert unmet_requirements(SimpleNamespace(G=None, n_phases=2, fluid=fl)) == ["model.fluid has no .vapoil"]
    A = sp.random(5, 5, density=0.5, random_state=0) + sp.eye(5)
    p = LinearizedProblemAD(A=A, b=np.arange(5.0), var_names=["p", "s"], var_sizes=[2, 3])
    M, rhs = problem_system(p)
    assert np.array_equal(rhs, -np.arange(5.0)) and np.array_equal(M.toarray(), A.toarray())
    assert [x.tolist() for x in split_increment(p, np.arange(5.0))] == [[0.0, 1.0], [2.0, 3.0, 4.0]]
    eq = SimpleNamespace(val=np.ones(2), jac=[sp.eye(2, format="csr"), sp.csr_array((2, 1))])
    eq2 = SimpleNamespace(val=np.zeros(1), jac=[sp.csr_array((1, 2)), sp.eye(1, format="csr")])
    q = LinearizedProblemAD.from_equations([eq, eq2], ["cell", "cell"], ["a", "b"], ["x", "y"])
    M, rhs = problem_system(q)
    assert M.shape == (3, 3) and np.array_equal(rhs, [-1.0, -1.0, 0.0])
    assert [x.size for x in split_increment(q, np.zeros(3))] == [2, 1]
