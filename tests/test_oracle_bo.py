"""The black-oil oracle against MRST's per-face flux fixture, central finite
differences of its own residual, and conservation of gas and water."""
import numpy as np
import pytest

import oracle.grid as og
from oracle.bo import BOFluid, BOModel, BOState, phase_flux, pvt_values, synthetic_deck
from oracle.ow import Ops

from cases import golden


def test_phase_flux_matches_mrst_fixture():
    """getFluxAndProps{Water,Oil,Gas}_BO on MRST's 3-cell, 2-face case:
    dp = grad(p_phase) - avg(rho) gdz, flag dp <= 0, kr and mu upwinded
    separately."""
    d = golden("bo_phase_flux.npz")
    ops = Ops(np.array([[0, 1], [1, 2]]), d["T"], 3)
    for ph, P in (("W", "pW"), ("O", "pO"), ("G", "pG")):
        kr = d["kr" + ph.lower()]
        v, up, dp = phase_flux(ops, d[P], d["rho" + ph], kr, d["mu" + ph], d["T"], d["gdz"])
        np.testing.assert_allclose(dp, d["dp" + ph], rtol=1e-13, atol=0)
        np.testing.assert_array_equal(up, d["up" + ph].astype(bool))
        np.testing.assert_allclose(v, d["v" + ph], rtol=1e-13, atol=0)


def test_pvt_parallel_interpolation_closed_forms():
    """On the saturated curve the parallel scheme returns the saturated
    table; halfway between two keys at a fixed offset above their bubble
    points it returns the average of the two branches' values there."""
    f = BOFluid(synthetic_deck())
    tab = f.bo_tab
    p = tab.sat_x[3:6]
    v, _, _ = pvt_values(tab, p, tab.key[3:6], np.ones(3, bool))
    np.testing.assert_allclose(v, tab.sat_f[3:6], rtol=1e-14)
    rs = 0.5 * (tab.key[4] + tab.key[5])
    ps = 0.5 * (tab.sat_x[4] + tab.sat_x[5])
    v, dfdp, _ = pvt_values(tab, np.array([ps + 20e5]), np.array([rs]), np.zeros(1, bool))
    b4 = tab.branches[4][1][0] + (tab.branches[4][1][1] - tab.branches[4][1][0]) * 20e5 / 50e5
    b5 = tab.branches[5][1][0] + (tab.branches[5][1][1] - tab.branches[5][1][0]) * 20e5 / 50e5
    np.testing.assert_allclose(v[0], 0.5 * (b4 + b5), rtol=1e-13)


def small_bo(nx=3, ny=3, nz=2, wells=True, gravity=og.NORM_GRAVITY, seed=0):
    g = og.cart_grid([nx, ny, nz], [30.0 * nx, 30.0 * ny, 5.0 * nz])
    rng = np.random.default_rng(seed)
    perm = 100 * og.MILLI * og.DARCY * np.exp(0.5 * rng.standard_normal(g.num_cells))
    rock = og.make_rock(g, perm, 0.2)
    t_half = og.compute_trans(g, rock)
    N, T = og.internal_connections(g, t_half)[:2]
    ws = []
    if wells:
        ws = og.add_well(None, g, rock, [0], "rate", 5e-4, compi=[0.0, 0.0, 1.0])
        ws = og.add_well(ws, g, rock, [g.num_cells - 1], "bhp", 150 * og.BARSA, compi=[0.0, 1.0, 0.0])
    fluid = BOFluid(synthetic_deck())
    return BOModel(g, rock, fluid, ws, N, T, gravity=gravity)


def _mixed_state(m, seed=1, gas=None):
    """Random pressures and saturations, free gas in `gas` cells (random
    half by default), away from the relperm table nodes."""
    rng = np.random.default_rng(seed)
    st = m.initial_state(200 * og.BARSA, 0.25, 0.0)
    nc = m.nc
    st.pressure = st.pressure + rng.normal(0, 5e5, nc)
    if gas is None:
        gas = rng.random(nc) < 0.5
    sg = np.where(gas, 0.1 + 0.1 * rng.random(nc), 0.0)
    sw = 0.25 + 0.1 * rng.random(nc)
    st.s = np.column_stack([sw, 1 - sw - sg, sg])
    rsat = m.fluid.rs_sat(st.pressure)
    st.rs = np.where(gas, rsat, 0.8 * rsat)
    st.status = np.where(gas, 3, 1)
    if m.wells:
        st.qGs[:] = [5e-4, 0.0]
        st.qOs[:] = [0.0, -2e-4]
        st.bhp[:] = [210 * og.BARSA, 150 * og.BARSA]
    return st


@pytest.mark.parametrize("wells", [False, True])
def test_jacobian_matches_finite_differences(wells):
    m = small_bo(wells=wells)
    st0 = _mixed_state(m, 2)
    st = _mixed_state(m, 3, gas=st0.status == 3)
    dt = 5 * og.DAY
    r, J, _ = m.residual_and_jacobian(st0, st, dt)
    J = J.toarray()
    x0 = np.concatenate([st.pressure, st.s[:, 0], np.where(st.status == 1, st.rs, st.s[:, 2]), st.qWs, st.qOs,
                         st.qGs, st.bhp])
    nc, nw = m.nc, len(m.wells)

    def resid(x):
        s = st.copy()
        s.pressure = x[:nc]
        sw, xv = x[nc:2 * nc], x[2 * nc:3 * nc]
        sg = np.where(s.status == 3, xv, 0.0)
        s.s = np.column_stack([sw, 1 - sw - sg, sg])
        s.rs = np.where(s.status == 1, xv, m.fluid.rs_sat(s.pressure))
        o = 3 * nc
        s.qWs, s.qOs, s.qGs, s.bhp = x[o:o + nw], x[o + nw:o + 2 * nw], x[o + 2 * nw:o + 3 * nw], x[o + 3 * nw:]
        return m.residual_and_jacobian(st0, s, dt)[0]

    scale = np.concatenate([np.full(nc, 1e3), np.full(2 * nc, 1e-6), np.full(3 * nw, 1e-9), np.full(nw, 1e3)])
    scale[2 * nc:3 * nc] = np.where(st.status == 1, 1e-4, 1e-6)
    for k in range(x0.size):
        h = scale[k]
        xp, xm = x0.copy(), x0.copy()
        xp[k] += h
        xm[k] -= h
        fd = (resid(xp) - resid(xm)) / (2 * h)
        colscale = np.abs(J[:, k]).max() + 1e-300
        assert np.abs(fd - J[:, k]).max() <= 1e-5 * colscale, f"column {k}"


def test_closed_box_conserves_water_and_gas():
    """No wells, gravity segregation of a mixed state: after a converged step
    the surface volumes of water and of gas (free + dissolved) are unchanged."""
    m = small_bo(wells=False)
    st0 = _mixed_state(m, 5)
    st1, ok, its = m.step(st0, 10 * og.DAY)
    assert ok

    def totals(st):
        f, p = m.fluid, st.pressure
        pv = m.pv_ref * f.pv_mult(p)
        bo = f.b_o(p, st.rs, st.status == 3)
        return (np.sum(pv * f.b_w(p) * st.s[:, 0]),
                np.sum(pv * (f.b_g(p) * st.s[:, 2] + st.rs * bo * st.s[:, 1])))

    w0, g0 = totals(st0)
    w1, g1 = totals(st1)
    assert abs(w1 - w0) <= 1e-7 * w0
    assert abs(g1 - g0) <= 1e-7 * g0
