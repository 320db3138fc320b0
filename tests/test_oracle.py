"""Pinning the CPU oracle against the reference's golden vectors, closed
forms and finite differences (no GPU)."""
import numpy as np
import pytest

import oracle.grid as og
from oracle.adi import ADI, ad_exp, init_variables_adi, value
from oracle.incomp import add_bc, add_source, incomp_tpfa
from oracle.ow import simulate_schedule

from cases import build_oracle, golden, p5_inputs, paired_inputs, small_gravity_inputs

MU1 = 1 * og.CENTI * og.POISE
RHO = np.array([1000.0])


def close(a, b, rtol, atol_floor=1e-15):
    a = np.ravel(np.asarray(a, float))
    b = np.ravel(np.asarray(b, float))
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() <= rtol * scale + atol_floor


def _mob(g):
    return np.full((g.num_cells, 1), 1.0 / MU1)


# -- computeTrans + incompTPFA vs MRST (examples/validation/validate_phase2.py) --
def test_incomp_qfs_matches_mrst():
    g = og.cart_grid([32, 32, 1], [500.0, 500.0, 10.0])
    rock = og.make_rock(g, (10 + (7 * np.arange(g.num_cells)) % 100) * og.MILLI * og.DARCY, 0.2)
    ht = og.compute_trans(g, rock)
    src = add_source(add_source(None, [0], 1e-3), [g.num_cells - 1], -1e-3)
    res = incomp_tpfa(g, ht, _mob(g), RHO, src=src)
    ref = golden("p2_qfs.npz")
    for key, val in (("hT", ht), ("p", res.pressure), ("flux", res.flux), ("fp", res.facePressure)):
        assert close(val, ref[key], 1e-9), key


def test_incomp_bc_matches_mrst():
    g = og.cart_grid([20, 10, 5], [200.0, 100.0, 50.0])
    rock = og.make_rock(g, (1 + (3 * np.arange(g.num_cells)) % 50) * og.MILLI * og.DARCY, 0.25)
    ht = og.compute_trans(g, rock)
    bd = np.any(g.neighbors == -1, axis=1)
    fc = g.face_centroids
    bc = add_bc(None, np.flatnonzero(bd & (fc[:, 0] < 1e-10)), "pressure", 200 * og.BARSA)
    bc = add_bc(bc, np.flatnonzero(bd & (fc[:, 0] > 200 - 1e-10)), "pressure", 100 * og.BARSA)
    bc = add_bc(bc, np.flatnonzero(bd & (fc[:, 1] < 1e-10)), "flux", 1e-6)
    res = incomp_tpfa(g, ht, _mob(g), RHO, bc=bc)
    ref = golden("p2_bc.npz")
    for key, val in (("hT", ht), ("p", res.pressure), ("flux", res.flux), ("fp", res.facePressure)):
        assert close(val, ref[key], 1e-9), key


def test_incomp_wells_gravity_matches_mrst():
    g = og.cart_grid([10, 10, 3], [1000.0, 1000.0, 30.0])
    rock = og.make_rock(g, np.tile(np.array([100.0, 100.0, 10.0]) * og.MILLI * og.DARCY, (g.num_cells, 1)), 0.3)
    ht = og.compute_trans(g, rock)
    gv = np.array([0, 0, og.NORM_GRAVITY])
    w = og.add_well(None, g, rock, [0, 100, 200], "rate", 5e-3, compi=[1.0], gravity=gv)
    w = og.add_well(w, g, rock, [99, 199, 299], "bhp", 100 * og.BARSA, compi=[1.0], gravity=gv)
    res = incomp_tpfa(g, ht, _mob(g), RHO, wells=w, gravity=gv)
    ref = golden("p2_wells.npz")
    pairs = (("hT", ht), ("p", res.pressure), ("flux", res.flux), ("wi1", w[0].WI), ("wi2", w[1].WI),
             ("dz1", w[0].dZ), ("dz2", w[1].dZ), ("wflux1", res.well_flux[0]), ("wflux2", res.well_flux[1]),
             ("bhp1", res.well_pressure[0]), ("bhp2", res.well_pressure[1]))
    for key, val in pairs:
        assert close(val, ref[key], 1e-9), key


def test_incomp_hydrostatic_column_matches_mrst():
    g = og.cart_grid([1, 1, 30], [1.0, 1.0, 300.0])
    rock = og.make_rock(g, 100 * og.MILLI * og.DARCY, 0.2)
    ht = og.compute_trans(g, rock)
    bd = np.any(g.neighbors == -1, axis=1)
    top = np.flatnonzero(bd & (g.face_centroids[:, 2] < 1e-10))
    res = incomp_tpfa(g, ht, _mob(g), RHO, bc=add_bc(None, top, "pressure", 100 * og.BARSA),
                      gravity=np.array([0, 0, og.NORM_GRAVITY]))
    ref = golden("p2_gravity.npz")
    assert close(ht, ref["hT"], 1e-9)
    assert close(res.pressure, ref["p"], 1e-9)
    assert np.abs(res.flux - ref["flux"]).max() < 1e-15
    # closed form: p(z) = p_top + rho g z at the cell centres
    z = g.cell_centroids[:, 2]
    assert np.allclose(res.pressure, 100 * og.BARSA + 1000.0 * og.NORM_GRAVITY * z, rtol=1e-12)


def test_trans_closed_form():
    """compute_trans.py:60-77: T = 2kA/dx on a uniform grid, anisotropic values."""
    g = og.cart_grid([3, 3, 3], [30.0, 30.0, 30.0])
    k = 100 * og.MILLI * og.DARCY
    assert np.allclose(og.compute_trans(g, og.make_rock(g, k, 0.2)), k * 100.0 / 5.0, rtol=1e-12)
    g = og.cart_grid([2, 2, 2], [2.0, 2.0, 2.0])
    t = og.compute_trans(g, og.make_rock(g, np.array([[1.0, 2.0, 4.0]]), 0.2))
    tags = g.cell_tags
    assert np.allclose(t[tags <= 1], 2.0) and np.allclose(t[(tags == 2) | (tags == 3)], 4.0)
    assert np.allclose(t[tags >= 4], 8.0)


# -- forward-mode AD against central differences ----------------------------
def test_adi_ops_match_finite_differences():
    rng = np.random.default_rng(1)
    x0 = rng.random(5) + 0.5
    y0 = rng.random(5) + 0.5

    def f(x, y):
        return ad_exp(0.3 * x) * y / (x + 2.0) - y ** 2.0 + (x * y)[np.array([4, 3, 2, 1, 0])]

    x, y = init_variables_adi(x0, y0)
    F = f(x, y)
    h = 1e-6
    for blk, base in ((0, x0), (1, y0)):
        J = F.jac[blk].toarray()
        for k in range(5):
            e = np.zeros(5)
            e[k] = h
            args_p = (x0 + e, y0) if blk == 0 else (x0, y0 + e)
            args_m = (x0 - e, y0) if blk == 0 else (x0, y0 - e)
            fd = (value(f(*args_p)) - value(f(*args_m))) / (2 * h)
            assert np.allclose(J[:, k], fd, rtol=1e-6, atol=1e-8)


def test_ow_jacobian_matches_finite_differences():
    """The oracle's AD Jacobian (gravity, c_r, cross-flowing well) against
    central differences of its residual (utils/adi/adi.py:266-301 method)."""
    inp = small_gravity_inputs()
    m, s0 = build_oracle(inp)
    rng = np.random.default_rng(3)
    st = s0.copy()
    st.pressure = st.pressure + rng.normal(0, 2e5, st.pressure.size)
    st.sW = np.clip(st.sW + rng.normal(0, 0.05, st.sW.size), 0.01, 0.99)
    st.qWs = st.qWs + 1e-4
    st.qOs = st.qOs - 1e-4
    cdp = m.compute_cdp_per_well(s0)
    dt = 5 * og.DAY
    r, J, _ = m.residual_and_jacobian(s0, st, dt, cdp)
    nc, nw = m.nc, len(m.wells)
    x0 = np.concatenate([st.pressure, st.sW, st.qWs, st.qOs, st.bhp])
    scales = np.concatenate([np.full(nc, 1e5), np.full(nc, 1e-2), np.full(2 * nw, 1e-4), np.full(nw, 1e5)])

    def res(x):
        s = st.copy()
        s.pressure, s.sW = x[:nc], x[nc:2 * nc]
        s.qWs, s.qOs, s.bhp = x[2 * nc:2 * nc + nw], x[2 * nc + nw:2 * nc + 2 * nw], x[2 * nc + 2 * nw:]
        return m.residual_and_jacobian(s0, s, dt, cdp)[0]

    cols = rng.choice(x0.size, 40, replace=False)
    Jd = J.toarray()
    for k in cols:
        h = 1e-6 * scales[k]
        e = np.zeros_like(x0)
        e[k] = h
        fd = (res(x0 + e) - res(x0 - e)) / (2 * h)
        ref = np.abs(Jd[:, k]).max() + 1e-30
        assert np.abs(Jd[:, k] - fd).max() <= 1e-5 * ref, k


# -- oil-water Newton vs the reference's own runs ------------------------------
def test_ow_paired_waterflood_matches_reference_run():
    """The reference's 15x15x3 waterflood: states and per-step Newton counts."""
    m, s0 = build_oracle(paired_inputs())
    states, its = simulate_schedule(s0, m, paired_inputs()["dts"])
    ref = golden("paired_waterflood.npz")
    assert np.array_equal(its, ref["newton_iterations"])
    assert close(np.vstack([s.pressure for s in states]), ref["pressure"], 1e-12)
    assert close(np.vstack([s.sW for s in states]), ref["sw"], 1e-12)
    assert close(np.vstack([s.qWs for s in states]), ref["qws"], 1e-10)
    assert close(np.vstack([s.qOs for s in states]), ref["qos"], 1e-10)
    assert close(np.vstack([s.bhp for s in states]), ref["bhp"], 1e-12)


def test_ow_matches_mrst_waterflood():
    """MRST's 15x15 waterflood, rtol 5e-6 (validate_phase5.py:55-63)."""
    inp = p5_inputs()
    m, s0 = build_oracle(inp)
    states, _ = simulate_schedule(s0, m, inp["dts"])
    ref = golden("p5_ow.npz")
    assert close(m.wells[0].WI, ref["WI1"], 1e-12) and close(m.wells[1].WI, ref["WI2"], 1e-12)
    assert close(np.vstack([s.pressure for s in states]), ref["p_hist"], 5e-6)
    assert close(np.vstack([s.sW for s in states]), ref["sw_hist"], 5e-6)
    assert close(np.vstack([s.qWs for s in states]), ref["qWs"], 5e-6)
    assert close(np.vstack([s.qOs for s in states]), ref["qOs"], 5e-6)
    assert close(np.vstack([s.bhp for s in states]), ref["bhp"], 5e-6)


def test_golden_run_detects_a_planted_upwind_mistake():
    """Downstream weighting planted in the oracle must fail the reference's
    golden run within two report steps (the check is not vacuous)."""
    import oracle.ow as oow
    inp = paired_inputs()
    m, s0 = build_oracle(inp)
    orig = oow.Ops.upw

    def wrong(self, flag, x):
        return orig(self, ~np.asarray(flag), x)

    oow.Ops.upw = wrong
    try:
        states, its = simulate_schedule(s0, m, inp["dts"][:2])
    except AssertionError:
        return   # Newton fails outright: detected
    finally:
        oow.Ops.upw = orig
    ref = golden("paired_waterflood.npz")
    p = np.vstack([s.pressure for s in states])
    assert not close(p, ref["pressure"][:2], 1e-6) or not np.array_equal(its, ref["newton_iterations"][:2])


def test_product_grid_setup_matches_oracle():
    """The product's host-side grid, transmissibility and well builders agree
    with the oracle's restatement (and thus with MRST's hT)."""
    import pymrst_amd as pa
    for inp in (paired_inputs(), p5_inputs(), small_gravity_inputs()):
        mo, _ = build_oracle(inp)
        g = pa.cart_grid(inp["dims"], inp["phys"])
        rock = pa.make_rock(g, inp["perm"], inp["poro"])
        N, T, _, _ = pa.tpfa_connections(g, pa.compute_trans(g, rock))
        assert np.array_equal(N, mo.ops.N)
        assert np.allclose(T, mo.ops.T, rtol=1e-14, atol=0)
        assert np.allclose(pa.pore_volume(g, rock), mo.pv, rtol=1e-15)
        grav = np.array([0, 0, inp["gravity"]]) if inp["gravity"] else None
        wells = None
        for cells, kind, val, compi, sign in inp["wells"]:
            wells = pa.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi), sign=sign, gravity=grav)
        for a, b in zip(wells, mo.wells):
            assert np.allclose(a.WI, b.WI, rtol=1e-14) and np.array_equal(a.dZ, b.dZ)
