"""The reference's linear-solver seam on the device: CPRSolver.solve(jac,
rhs, model, state) on the oracle's variable-blocked Jacobians equals SciPy's
direct solve to 1e-7 relative (the reference's own CPR contract,
linear_solvers.py:639-649), FP64 and FP32 preconditioners, oil-water and
black-oil, cross-flowing wells included; a second Newton system with the
same pattern reuses the device system; a solve that cannot converge raises
RuntimeError (the reference's cut signal)."""
from types import SimpleNamespace

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle.grid as og
from cases import build_oracle, small_gravity_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _ow_system(seed):
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    st = so.copy()
    rng = np.random.default_rng(seed)
    st.pressure = st.pressure + rng.normal(0, 2e5, st.pressure.size)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, st.sW.size), 0.0, 1.0)
    w = mo.wells[2]
    st.pressure[w.cells] = st.bhp[2] + 3e5 * np.where(np.arange(w.cells.size) % 2 == 0, 1.0, -1.0)
    r, J, _ = mo.residual_and_jacobian(so, st, 10 * og.DAY, mo.compute_cdp_per_well(so))
    f = mo.fluid
    model = SimpleNamespace(G=SimpleNamespace(cells=SimpleNamespace(num=mo.nc)), n_phases=2,
                            fluid=SimpleNamespace(disgas=False, vapoil=False,
                                                  b=lambda p, rs, sat, rv, op: [f.b(0, p), f.b(1, p)]))
    return J, r, model, {"pressure": st.pressure}


def _bo_system():
    from oracle.bo import BOFluid, BOModel, synthetic_deck
    dims = (6, 5, 3)
    nx, ny, nz = dims
    g = og.cart_grid(list(dims), [30.0 * nx, 30.0 * ny, 5.0 * nz])
    rng = np.random.default_rng(0)
    perm = 100 * og.MILLI * og.DARCY * np.exp(0.5 * rng.standard_normal(g.num_cells))
    rock = og.make_rock(g, perm, 0.2 + 0.05 * rng.random(g.num_cells))
    N, T = og.internal_connections(g, og.compute_trans(g, rock))[:2]
    nc = g.num_cells
    ws = og.add_well(None, g, rock, [nx * ny * k for k in range(nz)], "bhp", 200 * og.BARSA, compi=[0.0, 0.0, 1.0])
    # one perforation for the rate well: a rate well cross-flowing between two
    # perforations makes the Jacobian nearly singular in the uniform (p, bhp)
    # shift, where direct solvers themselves disagree beyond 1e-7; the bhp
    # well still cross-flows (its mixture couples its perforated cells)
    ws = og.add_well(ws, g, rock, [nc - 1], "rate", -3e-4, compi=[0.0, 1.0, 0.0])
    m = BOModel(g, rock, BOFluid(synthetic_deck()), ws, N, T, gravity=og.NORM_GRAVITY)
    st0 = m.initial_state(200 * og.BARSA, 0.25, 0.0)
    st = st0.copy()
    st.pressure = st.pressure + rng.normal(0, 5e5, nc)
    gas = rng.random(nc) < 0.5
    sg = np.where(gas, 0.15, 0.0)
    st.s = np.column_stack([np.full(nc, 0.3), 0.7 - sg, sg])
    rsat = m.fluid.rs_sat(st.pressure)
    st.rs = np.where(gas, rsat, 0.8 * rsat)
    st.status = np.where(gas, 3, 1)
    w = m.wells[0]
    st.pressure[w.cells] = st.bhp[0] + 8e5 * np.where(np.arange(w.cells.size) % 2 == 0, 1.0, -1.0)
    r, J = m.residual_and_jacobian(st0, st, 10 * og.DAY)[:2]
    f = m.fluid
    model = SimpleNamespace(G=SimpleNamespace(cells=SimpleNamespace(num=nc)), n_phases=3,
                            fluid=SimpleNamespace(disgas=True, vapoil=False,
                                                  b=lambda p, rs, sat, rv, op: [f.b_w(p), f.b_o(p, rs, sat),
                                                                                f.b_g(p)]))
    return J, r, model, {"pressure": st.pressure, "rs": st.rs}


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("kind", ["oil-water", "black-oil"])
def test_reference_jacobian_solve_matches_direct(kind, prec):
    from pymrst_amd.ow import CPRSolver
    J, r, model, state = _ow_system(11) if kind == "oil-water" else _bo_system()
    solver = CPRSolver(rtol=1e-10, precision=prec)
    x = solver.solve(J, -r, model=model, state=state)
    x_ref = spla.spsolve(J.tocsc(), -r)
    err = float(np.abs(x - x_ref).max() / np.abs(x_ref).max())
    assert err <= 1e-7, f"{kind} FP{prec}: {err:.2e} ({solver.last_iterations} iterations)"
    assert solver.last_iterations > 0
    res = np.linalg.norm(J @ x + r) / np.linalg.norm(r)
    assert res <= 1e-9, res


def test_device_system_is_reused_while_the_pattern_fits_and_failure_raises():
    from pymrst_amd.ow import CPRSolver
    solver = CPRSolver(rtol=1e-10)
    J1, r1, model, st1 = _ow_system(11)
    J2, r2, _, st2 = _ow_system(12)
    solver.solve(J1, -r1, model=model, state=st1)
    first = solver._sys
    # new values on the same pattern: the same device system
    J1b = J1.copy()
    J1b.data = J1b.data * (1.0 + 0.01 * np.random.default_rng(0).random(J1b.data.size))
    x1b = solver.solve(J1b, -r1, model=model, state=st1)
    assert solver._sys is first
    assert np.abs(x1b - spla.spsolve(J1b.tocsc(), -r1)).max() <= 1e-7 * np.abs(x1b).max()
    # another state: the pattern joins the new blocks (if any) and the solve
    # is right either way; the joined system then serves both
    x2 = solver.solve(J2, -r2, model=model, state=st2)
    x_ref = spla.spsolve(J2.tocsc(), -r2)
    assert np.abs(x2 - x_ref).max() <= 1e-7 * np.abs(x_ref).max()
    joined = solver._sys
    solver.solve(J1, -r1, model=model, state=st1)
    assert solver._sys is joined
    tight = CPRSolver(rtol=1e-14, maxiter=1, restart=2, use_pressure_stage=False, local_stage="block_jacobi")
    with pytest.raises(RuntimeError, match="did not converge"):
        tight.solve(J1, -r1, model=model, state=st1)


def test_ad_oo_seam_elliptic_injection_and_direct_fallback():
    """solve_linear_problem(problem, model) -> (dx, result, report) on the
    oil-water system equals the direct solve, split per primary variable;
    an injected elliptic solver (a SciPy direct solve of the equilibrated
    pressure system, the reference's AMGCL slot) replaces the device AMG
    and still gives the direct solution; a model missing the CPR
    attributes warns and solves directly on the device, and no model
    solves directly without a warning."""
    import warnings
    import scipy.sparse as sp
    from pymrst_amd.linear import LinearizedProblemAD
    from pymrst_amd.ow import CPRSolver
    J, r, model, state = _ow_system(11)
    x_ref = spla.spsolve(J.tocsc(), -r)
    nc = model.G.cells.num
    nw = (J.shape[0] - 2 * nc) // 3
    prob = LinearizedProblemAD(A=J, b=r, var_names=["pressure", "sW", "qWs", "qOs", "bhp"],
                               var_sizes=[nc, nc, nw, nw, nw], state=state)
    dx, result, report = CPRSolver(rtol=1e-10).solve_linear_problem(prob, model)
    assert np.abs(result - x_ref).max() <= 1e-7 * np.abs(x_ref).max()
    assert [d.size for d in dx] == [nc, nc, nw, nw, nw] and np.array_equal(np.concatenate(dx), result)
    assert report["Converged"] and report["Iterations"] > 0 and report["Residual"] <= 1e-9

    class Elliptic:
        def __init__(self):
            self.shapes = []

        def solve_linear_system(self, A, b):
            self.shapes.append(A.shape)
            d = A.diagonal()
            off = (A - sp.diags_array(d)).tocoo()
            assert (d > 0).all() and (off.data <= 1e-12).all()      # W J P is an M-matrix
            return spla.spsolve(sp.csc_matrix(A), b)

    ell = Elliptic()
    solver = CPRSolver(rtol=1e-10, elliptic_solver=ell)
    x = solver.solve(J, -r, model=model, state=state)
    assert np.abs(x - x_ref).max() <= 1e-7 * np.abs(x_ref).max()
    assert len(ell.shapes) >= solver.last_iterations >= 1 and set(ell.shapes) == {(nc, nc)}
    with pytest.warns(RuntimeWarning, match="fell back to a direct solve"):
        xf = CPRSolver().solve(J, -r, model=SimpleNamespace(G=model.G), state=state)
    assert np.abs(xf - x_ref).max() <= 1e-10 * np.abs(x_ref).max()
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        xn = CPRSolver().solve(J, -r)
        xs = CPRSolver().solve_linear_system(J, -r)
    assert np.abs(xn - x_ref).max() <= 1e-10 * np.abs(x_ref).max() and np.array_equal(xn, xs)
