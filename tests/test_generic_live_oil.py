"""GenericBlackOilModelAD's facility wells, three-phase live-oil form (CPU):
the oracle restatement (oracle/generic.py GenericBOModel) pinned against
the reference's own known answers, closed forms and invariants, and the
product's host-side facility helper for three phases (control_form3)
checked against the oracle.

Known answers and invariants:
* the reference's own proof (generic_black_oil_model.py:1116-1190): the
  generic model reproduces ThreePhaseBlackOilModelAD (SPE 1 there, the
  synthetic deck here, both with gravity and single-perforation wells at
  their reference depth) at tight tolerances, flash statuses included;
* a producing perforation carries the dissolved gas of its oil stream
  (facility_model.py:2466-2475): with no free gas the surface gas-oil ratio
  is the cell's rs;
* the surface stream an injecting perforation delivers, taken back to
  reservoir conditions with the dissolution split (_volume_ratio,
  facility_model.py:2391-2432: (mix_g - rs mix_o) / b_g for the gas), fills
  exactly the perforation's reservoir volume rate;
* the wellbore column of a producer without history in oil-only cells
  weighs live oil, b_o (rhoS_o + rs rhoS_g) g dZ, dissolved gas included
  (connection_pressure_drop :2294-2389 with Density);
* rhoScale = rhoS / mean(rho) over the whole grid at state0, with the oil
  density b_o (rhoS_o + rs rhoS_g) (facility_model.py:2239-2263: "oil's
  true density is b_O*(rhoOS + rs*rhoGS)");
* update_state_black_oil_generic (update_state_black_oil_generic.py:30-80):
  the saturation changes are scaled together so the largest is ds_max
  (hand-computed), the pressure change is capped at dp_max_rel |p|;
* over a converged step with gravity, multi-perforation wells, a shut well
  and cross-flow, water, oil and gas (free plus dissolved) in place change
  by dt times the wells' surface rates;
* the facility Jacobian equals central differences of its residual;
* a rate injector asked for more than its bhp limit buys switches to bhp
  and holds it (the reference's switching self-test, three phases); a shut
  well has zero rates and the mean perforation pressure as bhp.
"""
import numpy as np
import pytest

import oracle.grid as og
from oracle.adi import value
from oracle.bo import BOFluid, BOModel, synthetic_deck
from oracle.generic import GenericBOModel, control_equation
from oracle.nonlinear import StabilisedNewton

from mutation import fails, planted

BARSA, DAY, MD = og.BARSA, og.DAY, og.MILLI * og.DARCY


def _grid(nx=4, ny=4, nz=5, seed=2):
    g = og.cart_grid([nx, ny, nz], [20.0 * nx, 20.0 * ny, 5.0 * nz])
    rng = np.random.default_rng(seed)
    rock = og.make_rock(g, np.exp(rng.normal(np.log(200 * MD), 0.5, (g.num_cells, 1))), np.full(g.num_cells, 0.2))
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    return g, rock, N, T


def _live_state(m, g, cap_layers=1, p_top=200 * BARSA):
    """Hydrostatic pressure, a gas cap (status 3) in the top layers over
    undersaturated oil (status 1, rs 85 % of saturation), sW 0.25."""
    st = m.initial_state(p_top, 0.25, 0.0)
    z = g.cell_centroids[:, 2]
    st.pressure = p_top + 800.0 * og.NORM_GRAVITY * (z - z.min())
    nx, ny = g.dims[0], g.dims[1]
    cap = np.arange(m.nc) < cap_layers * nx * ny
    sg = np.where(cap, 0.15, 0.0)
    st.s = np.column_stack([np.full(m.nc, 0.25), 0.75 - sg, sg])
    rsat = value(m.fluid.rs_sat(st.pressure))
    st.rs = np.where(cap, rsat, 0.85 * rsat)
    st.status = np.where(cap, 3, 1).astype(np.int64)
    return st


def _gravity_case(mod=None, shut=True, lims=None):
    """4x4x5 with gravity: a gas injector perforated in the two top layers
    (a rate well), a bhp producer perforated in four layers (its top
    perforation takes gas from the cap), a shut producer."""
    M = GenericBOModel if mod is None else mod.GenericBOModel
    g, rock, N, T = _grid()
    grav = np.array([0, 0, og.NORM_GRAVITY])
    col = lambda i, j, ks: np.array([i + 4 * (j + 4 * k) for k in ks])
    wells = og.add_well(None, g, rock, col(0, 0, [0, 1]), "rate", 0.05, compi=np.array([0.0, 0.0, 1.0]), sign=1,
                        gravity=grav, lims=lims)
    wells = og.add_well(wells, g, rock, col(3, 3, [0, 1, 2, 3]), "bhp", 195 * BARSA,
                        compi=np.array([0.0, 1.0, 0.0]), sign=-1, gravity=grav)
    if shut:
        wells = og.add_well(wells, g, rock, col(0, 3, [2, 3]), "bhp", 190 * BARSA, compi=np.array([0.0, 1.0, 0.0]),
                            sign=-1, gravity=grav, status=False)
    m = M(g, rock, BOFluid(synthetic_deck()), wells, N, T, gravity=og.NORM_GRAVITY)
    st = _live_state(m, g)
    st.bhp[0] = 215 * BARSA
    st.qGs[0], st.qOs[0], st.qWs[0] = 0.05, 0.0, 0.0
    return m, st


# -- the reference's own proof, on the synthetic deck -------------------------
def test_live_oil_generic_reproduces_the_hand_rolled_model():
    g, rock, N, T = _grid(seed=5)
    fluid = BOFluid(synthetic_deck())
    grav = np.array([0, 0, og.NORM_GRAVITY])
    depth = g.cell_centroids[:, 2]
    inj, prod = 5, g.num_cells - 3
    wells = og.add_well(None, g, rock, [inj], "rate", 0.02, compi=np.array([0.0, 0.0, 1.0]), sign=1,
                        gravity=grav, ref_depth=float(depth[inj]))
    wells = og.add_well(wells, g, rock, [prod], "bhp", 190 * BARSA, compi=np.array([0.0, 1.0, 0.0]), sign=-1,
                        gravity=grav, ref_depth=float(depth[prod]))
    tight = dict(tol_cnv=1e-8, tol_mb=1e-12, tol_well=1e-8)
    gen = GenericBOModel(g, rock, fluid, wells, N, T, gravity=og.NORM_GRAVITY, **tight)
    hand = BOModel(g, rock, fluid, wells, N, T, gravity=og.NORM_GRAVITY, **tight)
    s_g = _live_state(gen, g)
    s_h = s_g.copy()
    s_h = hand.initial_state(200 * BARSA, 0.25, 0.0)
    for name in ("pressure", "s", "rs", "status", "qWs", "qOs", "qGs", "bhp"):
        setattr(s_h, name, np.array(getattr(s_g, name), copy=True))
    newton = StabilisedNewton(max_iterations=25)
    worst = 0.0
    for _ in range(3):
        s_g, ok_g, _, _ = newton.step(gen, s_g, 2 * DAY)
        s_h, ok_h, _ = hand.step(s_h, 2 * DAY)
        assert ok_g and ok_h
        assert np.array_equal(s_g.status, s_h.status)
        for a, b in ((s_g.pressure, s_h.pressure), (s_g.s, s_h.s), (s_g.rs, s_h.rs), (s_g.qOs, s_h.qOs),
                     (s_g.qGs, s_h.qGs), (s_g.bhp, s_h.bhp)):
            worst = max(worst, np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
    assert worst < 1e-10, worst


# -- perforation closed forms -----------------------------------------------
def _one_cell_model(status, rs_frac, sw, sg, bhp, compi):
    g, rock, N, T = _grid(2, 1, 1)
    wells = og.add_well(None, g, rock, [0], "bhp", bhp, compi=np.array(compi), sign=-1)
    m = GenericBOModel(g, rock, BOFluid(synthetic_deck()), wells, N, T)
    st = m.initial_state(200 * BARSA, sw, sg)
    st.status[:] = status
    st.rs = rs_frac * value(m.fluid.rs_sat(st.pressure))
    st.s = np.column_stack([np.full(2, sw), np.full(2, 1 - sw - sg), np.full(2, sg)])
    st.bhp[:] = bhp
    return m, st


def _perforation_rates(m, st):
    """Surface rates (water, oil, gas) of well 0's single perforation and
    the cell's b and rs at st."""
    f = m.fluid
    p = st.pressure
    st1, st3 = st.status == 1, st.status == 3
    rs = np.where(st1, st.rs, value(f.rs_sat(p)))
    sw, sg = st.s[:, 0], np.where(st3, st.s[:, 2], 0.0)
    b = (value(f.b_w(p)), value(f.b_o(p, rs, st3)), value(f.b_g(p)))
    mu = (value(f.mu_w(p)), value(f.mu_o(p, rs, st3)), value(f.mu_g(p)))
    kr = (value(f.kr_w(sw)), value(f.kr_o(sw, 1 - sw - sg, sg)), value(f.kr_g(sg)))
    mob = tuple(kr[i] / mu[i] for i in range(3))
    m._cdp = [np.zeros(w.cells.size) for w in m.wells]
    m._flux = [None] * len(m.wells)
    q = (st.qWs[:1], st.qOs[:1], st.qGs[:1])
    cq, _, _ = m._well(0, p, mob, b, rs, q, st.bhp[:1])
    return [float(np.ravel(value(c))[0]) for c in cq], [x[0] for x in b], rs[0], m._flux[0][0]


def test_gas_oil_ratio_of_an_undersaturated_producer_is_rs():
    m, st = _one_cell_model(1, 0.8, 0.2, 0.0, 150 * BARSA, [0.0, 1.0, 0.0])
    (qw, qo, qg), _, rs, _ = _perforation_rates(m, st)
    assert qo < 0 and qw == pytest.approx(0.0, abs=1e-30)
    assert qg == pytest.approx(rs * qo, rel=1e-13)


def test_injected_surface_stream_fills_the_perforation_volume():
    m, st = _one_cell_model(1, 0.7, 0.3, 0.0, 260 * BARSA, [0.0, 0.3, 0.7])
    (qw, qo, qg), (bw, bo, bg), rs, vol = _perforation_rates(m, st)
    reservoir = qw / bw + qo / bo + (qg - rs * qo) / bg
    assert qg > 0 and qo > 0
    assert reservoir == pytest.approx(vol.sum(), rel=1e-13)
    assert qo / qg == pytest.approx(0.3 / 0.7, rel=1e-13)
    # the perforation's reservoir flux handed to the next column is the
    # total split by compi (cq_vol, facility_model.py:2511-2528)
    np.testing.assert_allclose(vol, vol.sum() * np.array([0.0, 0.3, 0.7]), rtol=1e-13)


def test_column_of_live_oil_weighs_its_dissolved_gas():
    g, rock, N, T = _grid(2, 1, 3)
    grav = np.array([0, 0, og.NORM_GRAVITY])
    cell = [4]                                      # bottom layer
    wells = og.add_well(None, g, rock, cell, "bhp", 150 * BARSA, compi=np.array([0.0, 1.0, 0.0]), sign=-1,
                        gravity=grav)
    m = GenericBOModel(g, rock, BOFluid(synthetic_deck()), wells, N, T, gravity=og.NORM_GRAVITY)
    st = m.initial_state(220 * BARSA, 0.2, 0.0)     # sW at connate: only oil flows
    st.rs = 0.9 * value(m.fluid.rs_sat(st.pressure))
    cdp = m.compute_cdp_per_well(st)[0]
    f = m.fluid
    p, rs = st.pressure[cell], st.rs[cell]
    rho_o = value(f.b_o(p, rs, np.array([False]))) * (f.rhoS[1] + rs * f.rhoS[2])
    assert cdp[0] == pytest.approx(og.NORM_GRAVITY * float(rho_o[0]) * float(wells[0].dZ[0]), rel=1e-13)
    assert f.rhoS[2] * rs[0] > 0.01 * f.rhoS[1]     # the dissolved gas is not negligible here


def test_rho_scale_takes_the_live_oil_density_over_the_grid():
    m, st = _gravity_case()
    m.compute_cdp_per_well(st)
    f = m.fluid
    p, rs = st.pressure, np.where(st.status == 1, st.rs, value(f.rs_sat(st.pressure)))
    b_o = value(f.b_o(p, rs, st.status == 3))
    oil_mass_per_volume = b_o * f.rhoS[1] + b_o * rs * f.rhoS[2]     # surface oil + dissolved gas, per rm3
    assert m._rho_scale[1] == pytest.approx(f.rhoS[1] / oil_mass_per_volume.mean(), rel=1e-13)
    assert m._rho_scale[2] == pytest.approx(1.0 / value(f.b_g(p)).mean(), rel=1e-13)


# -- the generic update ---------------------------------------------------------
def test_generic_update_scales_saturation_changes_together():
    m, st = _one_cell_model(3, 1.0, 0.3, 0.2, 150 * BARSA, [0.0, 1.0, 0.0])
    nc, nw = m.nc, 1
    upd = np.zeros(3 * nc + 4 * nw)
    upd[0] = 0.5 * st.pressure[0]                 # +50 %: capped at +20 %
    upd[nc] = 0.5                                # dsw
    upd[2 * nc] = -0.2                           # dsg: ds = (0.5, -0.3, -0.2), largest 0.5
    p0 = st.pressure[0]
    m.apply_update(st, upd)
    assert st.pressure[0] == pytest.approx(1.2 * p0, rel=1e-15)
    np.testing.assert_allclose(st.s[0], [0.3 + 0.2, 0.5 - 0.12, 0.2 - 0.08], rtol=1e-14)
    assert st.status[0] == 3


# -- converged steps ------------------------------------------------------------
def _in_place(m, st):
    f, p = m.fluid, st.pressure
    pv = m.pv_ref * value(f.pv_mult(p))
    st3 = st.status == 3
    rs = np.where(st.status == 1, st.rs, value(f.rs_sat(p)))
    so = st.s[:, 1]
    sg = np.where(st3, st.s[:, 2], 0.0)
    bo = value(f.b_o(p, rs, st3))
    return (np.sum(pv * value(f.b_w(p)) * st.s[:, 0]), np.sum(pv * bo * so),
            np.sum(pv * (value(f.b_g(p)) * sg + rs * bo * so)))


def test_a_converged_live_oil_step_balances_the_wells():
    m, s0 = _gravity_case()
    m.tol_cnv, m.tol_mb, m.tol_well = 1e-9, 1e-12, 1e-9
    dt = 2 * DAY
    st, ok, its, _ = StabilisedNewton().step(m, s0, dt)
    assert ok
    before, after = _in_place(m, s0), _in_place(m, st)
    for i, q in enumerate((st.qWs, st.qOs, st.qGs)):
        assert after[i] - before[i] == pytest.approx(dt * q.sum(), rel=1e-6, abs=1e-9 * before[i])
    # the shut well: no flow, bhp at the mean of its perforations' pressures
    assert st.qWs[2] == st.qOs[2] == st.qGs[2] == 0.0
    assert st.bhp[2] == pytest.approx(st.pressure[m.wells[2].cells].mean(), rel=1e-12)
    # the producer's top perforation sees the gas cap: its gas rate exceeds rs x oil
    assert st.qGs[1] < 0 and abs(st.qGs[1]) > abs(st.qOs[1]) * st.rs[m.wells[1].cells[-1]]


def test_live_oil_facility_jacobian_equals_central_differences():
    m, s0 = _gravity_case()
    st = s0.copy()
    rng = np.random.default_rng(8)
    st.pressure = st.pressure + rng.normal(0, 1e5, m.nc)
    # saturations off the relative-permeability table nodes (kinks there)
    sw = 0.25 + 0.02 * rng.random(m.nc)
    sg = np.where(st.status == 3, 0.13 + 0.04 * rng.random(m.nc), 0.0)
    st.s = np.column_stack([sw, 1 - sw - sg, sg])
    st.bhp = np.array([215 * BARSA, 202 * BARSA, 190 * BARSA])   # the producer's top perforation injects
    st.qWs = np.array([0.0, -1e-5, 0.0])
    st.qOs = np.array([0.0, -2e-4, 0.0])
    st.qGs = np.array([0.05, -0.01, 0.0])
    dt = 3 * DAY
    cdp = m.compute_cdp_per_well(s0)
    r0, J, _ = m.residual_and_jacobian(s0, st, dt, cdp)
    nc, nw = m.nc, 3
    xv = np.where(st.status == 1, st.rs, st.s[:, 2])
    x0 = np.concatenate([st.pressure, st.s[:, 0], xv, st.qWs, st.qOs, st.qGs, st.bhp])
    Jd = J.toarray()

    def res(x):
        s = st.copy()
        s.pressure = x[:nc].copy()
        sw, x3 = x[nc:2 * nc], x[2 * nc:3 * nc]
        sg = np.where(s.status == 3, x3, 0.0)
        s.s = np.column_stack([sw, 1 - sw - sg, sg])
        s.rs = np.where(s.status == 1, x3, s.rs)
        o = 3 * nc
        s.qWs, s.qOs, s.qGs, s.bhp = (x[o:o + nw].copy(), x[o + nw:o + 2 * nw].copy(),
                                      x[o + 2 * nw:o + 3 * nw].copy(), x[o + 3 * nw:].copy())
        return m.residual_and_jacobian(s0, s, dt, cdp)[0]

    cols = list(rng.choice(3 * nc, 15, replace=False)) + list(range(3 * nc, 3 * nc + 4 * nw))
    for c in cols:
        h = (1e-3 if c < nc else 1e-7) if c < 3 * nc else (1e-2 if c >= 3 * nc + 3 * nw else 1e-9)
        if nc * 2 <= c < 3 * nc and st.status[c - 2 * nc] == 1:
            h = 1e-4
        xp, xm = x0.copy(), x0.copy()
        xp[c] += h
        xm[c] -= h
        fd = (res(xp) - res(xm)) / (2 * h)
        scale = np.maximum(np.abs(Jd[:, c]), np.abs(fd)).max()
        assert np.abs(Jd[:, c] - fd).max() <= 1e-5 * max(scale, 1e-30), c


def test_gas_injector_switches_from_rate_to_bhp():
    m, s0 = _gravity_case(shut=False, lims={"bhp": 205 * BARSA})
    m.wells[0].val = 0.5                           # far more than 205 bar buys
    s0.qGs[0] = 0.5
    st, ok, its, _ = StabilisedNewton().step(m, s0, DAY)
    assert ok
    assert m.wells[0].type == "bhp" and m.switches[0][:3] == (0, "rate", "bhp")
    assert st.bhp[0] == pytest.approx(205 * BARSA, rel=1e-10)
    assert 0 < st.qGs[0] < 0.5


# -- the product's control rows -------------------------------------------------
def test_control_rows_oracle_and_product_three_phases():
    from pymrst_amd.generic import control_form3
    from pymrst_amd.grid import Well

    q = (np.array([-1.0]), np.array([-2.0]), np.array([-30.0]))
    cases = [("bhp", 1.5e7, -1), ("rate", -10.0, -1), ("wrat", -1.0, -1), ("orat", -2.5, -1), ("grat", -40.0, -1),
             ("lrat", -3.5, -1), ("orat", 0.0, -1), ("grat", 0.0, -1), ("bhp", 2e7, 1)]
    for kind, val, sign in cases:
        ow = og.Well(cells=np.array([0]), type=kind, val=val, WI=np.ones(1), dZ=np.zeros(1),
                     compi=np.array([0.0, 1.0, 0.0]), sign=sign)
        pw = Well(cells=np.array([0]), type=kind, val=val, WI=np.ones(1), dZ=np.zeros(1),
                  compi=np.array([0.0, 1.0, 0.0]), sign=sign)
        bhp = np.array([1.2e7])
        ref = float(np.ravel(value(control_equation(ow, list(q), bhp)))[0])
        coef, target = control_form3(pw, -1.0, -2.0, -30.0)
        got = coef[0] * -1.0 + coef[1] * -2.0 + coef[2] * -30.0 + coef[3] * 1.2e7 - target
        assert got == pytest.approx(ref, rel=1e-14, abs=1e-9), (kind, val, sign)


# -- planted mistakes -------------------------------------------------------------
def test_planted_mistakes_are_caught():
    import oracle.generic as mod
    plants = [
        (mod.GenericBOModel, "_well", ("cq_ps[2] = cq_ps[2] + rs_c * (b_c[1] * cq_p[1])", "pass"),
         test_gas_oil_ratio_of_an_undersaturated_producer_is_rs),
        (mod.GenericBOModel, "_well", ("(mix[2] - rs_c * mix[1]) / b_c[2]", "(mix[2] + rs_c * mix[1]) / b_c[2]"),
         test_injected_surface_stream_fills_the_perforation_volume),
        (mod.GenericBOModel, "densities", ("b[1] * (f.rhoS[1] + rs * f.rhoS[2])", "b[1] * f.rhoS[1]"),
         test_column_of_live_oil_weighs_its_dissolved_gas),
        (mod.GenericBOModel, "densities", ("b[1] * (f.rhoS[1] + rs * f.rhoS[2])", "b[1] * f.rhoS[1]"),
         test_rho_scale_takes_the_live_oil_density_over_the_grid),
        (mod.GenericBOModel, "apply_update", ("st.s = prev.s + ds * scale[:, None]",
                                              "st.s = prev.s + np.clip(ds, -self.ds_max, self.ds_max)"),
         test_generic_update_scales_saturation_changes_together),
        (mod.GenericBOModel, "_well", ("self._flux[k] = vol", "self._flux[k] = vol * 0.0 + vol[:, ::-1]"),
         test_injected_surface_stream_fills_the_perforation_volume),
        (mod.GenericBOModel, "_well",
         ("ref = ad_sum(p[wc]) / max(wc.size, 1)", "ref = ad_sum(p[wc]) / max(wc.size + 1, 1)"),
         test_a_converged_live_oil_step_balances_the_wells),
    ]
    for obj, name, edit, check in plants:
        check()
        with planted(obj, name, edit):
            assert fails(check), f"planted mistake not caught: {edit[0]!r}"
