"""GenericBlackOilModelAD's facility wells, oil-water form (CPU): the
oracle restatement (oracle/generic.py) pinned against the reference's own
known answers, and the product's host-side facility helpers
(pymrst_amd/generic.py: limit_broken, control_form, column_drops) checked
against the oracle and the same answers.

Known answers:
* the generic model's self-tests (generic_black_oil_model.py:1015-1250):
  it converges on the 8x8x1 case; it reproduces the hand-rolled
  TwoPhaseOilWaterModelAD to 1e-11 at tight tolerances; a rate injector
  asked for more than its 250 bar limit buys switches to bhp and holds it;
  a limit never reached leaves the well alone; a bhp producer pulling past
  its liquid-rate limit switches to lrat and holds it exactly;
* the switching rules (tests/ad_core/models/test_well_limits_and_resv.py:
  80-165): an injector is limited from above and a producer from below, an
  unset limit is never broken, a well is never switched to the control it
  is on, the first broken limit in MRST's order wins;
* the control rows (same file :340-400 and facility_model.py:2600-2631): a
  zero rate target closes on the total rate, a nonzero target on its own
  phase, a zero bhp target is a bhp target, a bhp well flowing the wrong
  way holds its total rate at zero;
* the wellbore column (facility_model.py:2294-2389): a uniform stream
  weighs rho g dZ; fluid entering only at the top weighs the whole column
  with the top's density; an injector weighs its compi mixture; a producer
  without history weighs WI x mobility;
* rhoScale = rhoS / mean(rho) over the whole grid at the step's start
  (facility_model.py:2234-2266);
* over a converged step the water and oil in place change by dt times the
  wells' surface rates;
* the facility Jacobian equals central differences of its residual, with
  gravity, multi-perforation wells and an injecting perforation of a
  producer.
"""
import sys
import types

import numpy as np
import pytest

import oracle.grid as og
from oracle.generic import (GenericOWModel, GenericOWState, column_pressure_drop, control_equation,
                            violated_limit)
from oracle.ow import OWFluid, OWModel
from oracle.nonlinear import StabilisedNewton
from oracle.adi import ADI, init_variables_adi, value

from mutation import fails, planted

BARSA, DAY = og.BARSA, og.DAY
MD = 1e-3 * 9.869233e-13
CP = 1e-3


def _fluid(c=(1e-6 / BARSA, 1e-4 / BARSA), c_r=None):
    return OWFluid(mu=np.array([1.0, 5.0]) * CP, rhoS=np.array([1000.0, 800.0]), n=np.array([2.0, 2.0]),
                   c=np.asarray(c, float), p_ref=200 * BARSA, c_r=c_r)


def _ow_case(wspec, tight=False, generic=True, mod=None):
    """generic_black_oil_model.py:989-1000: 8x8x1 cells of 200x200x10 m,
    100 mD, porosity 0.25."""
    M = GenericOWModel if generic else OWModel
    if mod is not None:
        M = mod.GenericOWModel
    g = og.cart_grid([8, 8, 1], [200.0, 200.0, 10.0])
    rock = og.make_rock(g, np.full((64, 1), 100 * MD), np.full(64, 0.25))
    wells = None
    for cells, kind, val, compi, sign, lims in wspec:
        wells = og.add_well(wells, g, rock, np.array(cells), kind, val, compi=np.array(compi), sign=sign, lims=lims)
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    tol = dict(tol_cnv=1e-8, tol_mb=1e-12, tol_well=1e-10) if tight else {}
    m = M(g, rock, _fluid(), wells, N, T, **tol)
    return m, m.initial_state(200 * BARSA, 0.2)


_BASE = [([0], "rate", 1e-3, [1, 0], 1, None), ([63], "bhp", 150 * BARSA, [0, 1], -1, None)]


# -- the reference's generic-model self-tests ---------------------------------
def test_generic_model_converges():
    m, s = _ow_case(_BASE)
    st, ok, its, _ = StabilisedNewton().step(m, s, 30 * DAY)
    assert ok
    assert np.isclose(st.qWs[0] + st.qOs[0], 1e-3, rtol=1e-9)
    assert np.isclose(st.bhp[1], 150 * BARSA)
    assert st.well_flux is not None and len(st.well_flux) == 2


def test_generic_model_reproduces_the_hand_rolled_model():
    mg, sg = _ow_case(_BASE, tight=True)
    mh, sh = _ow_case(_BASE, tight=True, generic=False)
    worst = 0.0
    for _ in range(4):
        sg, okg, _, _ = StabilisedNewton().step(mg, sg, 30 * DAY)
        sh, okh, _ = mh.step(sh, 30 * DAY)
        assert okg and okh
        for a, b in ((sg.pressure, sh.pressure), (sg.sW, sh.sW), (sg.qWs, sh.qWs), (sg.qOs, sh.qOs),
                     (sg.bhp, sh.bhp)):
            worst = max(worst, np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
    assert worst < 1e-11, worst


def test_injector_switches_from_rate_to_bhp():
    p_lim = 250.0 * BARSA
    m, s = _ow_case([([0], "rate", 5e-3, [1, 0], 1, {"bhp": p_lim}), ([63], "bhp", 150 * BARSA, [0, 1], -1, None)])
    st, ok, _, _ = StabilisedNewton().step(m, s, 30 * DAY)
    assert ok and m.wells[0].type == "bhp"
    assert np.isclose(st.bhp[0], p_lim)
    assert 0 < st.qWs[0] < 5e-3


def test_well_within_its_limits_does_not_switch():
    m, s = _ow_case([([0], "rate", 1e-3, [1, 0], 1, {"bhp": 1000 * BARSA}),
                     ([63], "bhp", 150 * BARSA, [0, 1], -1, None)])
    st, ok, _, _ = StabilisedNewton().step(m, s, 30 * DAY)
    assert ok and m.wells[0].type == "rate" and not m.switches
    assert np.isclose(st.qWs[0] + st.qOs[0], 1e-3, rtol=1e-9)


def test_producer_switches_on_liquid_rate():
    l_lim = -5e-4
    m, s = _ow_case([([0], "rate", 2e-3, [1, 0], 1, None), ([63], "bhp", 100 * BARSA, [0, 1], -1, {"lrat": l_lim})])
    st, ok, _, _ = StabilisedNewton().step(m, s, 30 * DAY)
    assert ok and m.wells[1].type == "lrat"
    assert np.isclose(st.qWs[1] + st.qOs[1], l_lim, rtol=1e-8)
    assert st.bhp[1] > 100 * BARSA


# -- switching rules ---------------------------------------------------------------
def _w(sign, lims, kind="rate", val=-3.0, status=True):
    return types.SimpleNamespace(sign=sign, lims=lims, type=kind, val=val, status=status)


def _switch_rules(fn):
    assert fn(_w(-1, {"bhp": 2.5e7}), 2.0e7, [0.0, 0.0]) == "bhp"           # producer below its floor
    assert fn(_w(-1, {"bhp": 1.0e7}), 2.0e7, [0.0, 0.0]) is None            # above it
    assert fn(_w(1, {"bhp": 3.0e7}, "bhp"), 2.0e7, [1.0, 0.0]) is None      # unset rate limit not broken
    assert fn(_w(1, {"bhp": 2.5e7}), 3.0e7, [1.0, 0.0]) == "bhp"            # injector above its ceiling
    assert fn(_w(-1, {"bhp": 2.0e7}, "bhp"), 1.9e7, [0.0, 0.0]) is None     # never to its own control
    assert fn(_w(-1, {"bhp": 2.5e7, "orat": -1.0}), 2.0e7, [0.0, -5.0]) == "bhp"   # bhp before orat
    assert fn(_w(-1, {"orat": -1.0}, "bhp"), 2.0e7, [0.0, -5.0]) == "orat"
    assert fn(_w(-1, {"lrat": -1.0}, "bhp"), 2.0e7, [-4.0, -4.0]) == "lrat"
    assert fn(_w(-1, {}), 2.0e7, [0.0, 0.0]) is None
    assert fn(_w(1, {"rate": 4.0}, "bhp"), 1.0e6, [5.0, 5.0]) == "rate"


def test_switching_rules_oracle():
    _switch_rules(violated_limit)


def test_switching_rules_product():
    from pymrst_amd.generic import limit_broken
    _switch_rules(limit_broken)


# -- control rows ----------------------------------------------------------------
def _qs(qw, qo):
    return [np.array([qw]), np.array([qo])]


def test_control_rows_oracle_and_product():
    from pymrst_amd.generic import control_form
    cases = [  # well, (qw, qo, bhp), expected residual
        (_w(-1, None, "orat", 0.0), (-2.0, -3.0, 1e7), -5.0),          # zero target: total rate
        (_w(-1, None, "orat", -1.0), (-2.0, -3.0, 1e7), -2.0),         # own phase
        (_w(-1, None, "wrat", -1.0), (-2.0, -3.0, 1e7), -1.0),
        (_w(-1, None, "lrat", -1.0), (-2.0, -3.0, 1e7), -4.0),
        (_w(1, None, "rate", 4.0), (1.0, 2.0, 1e7), -1.0),
        (_w(-1, None, "bhp", 0.0), (-2.0, -3.0, 1e7), 1e7),            # zero bhp target stays bhp
        (_w(-1, None, "bhp", 1.5e7), (2.0, 3.0, 1e7), 5.0),           # producer flowing in: total rate
        (_w(-1, None, "bhp", 1.5e7), (-2.0, -3.0, 1e7), -5e6),
    ]
    for w, (qw, qo, bhp), want in cases:
        got = float(np.ravel(value(control_equation(w, _qs(qw, qo), np.array([bhp]))))[0])
        assert got == pytest.approx(want), (w.type, w.val, got, want)
        (cw, co, cb), target = control_form(w, qw, qo)
        assert cw * qw + co * qo + cb * bhp - target == pytest.approx(want)


# -- wellbore column -------------------------------------------------------------
def _column_impls():
    from pymrst_amd.generic import column_drops
    return (column_pressure_drop, column_drops)


@pytest.mark.parametrize("impl", [0, 1])
def test_column_uniform_stream_and_top_only_stream(impl):
    fn = _column_impls()[impl]
    g = 9.80665
    dz = np.array([2.0, 5.0, 9.0])
    wi = np.ones(3)
    rho = np.array([[1000.0, 800.0], [1010.0, 790.0], [1020.0, 780.0]])
    mob = np.full((3, 2), 0.5)
    # only the bottom perforation flows, pure water: every segment carries it
    flux = np.array([[0.0, 0.0], [0.0, 0.0], [-2.0, 0.0]])
    cdp = fn(dz, wi, rho, mob, [0, 1], -1, flux, g)
    np.testing.assert_allclose(cdp, g * 1020.0 * dz, rtol=1e-14)
    # only the top perforation flows, pure oil: the column weighs its oil
    flux = np.array([[0.0, -1.0], [0.0, 0.0], [0.0, 0.0]])
    cdp = fn(dz, wi, rho, mob, [0, 1], -1, flux, g)
    np.testing.assert_allclose(cdp, g * 800.0 * dz, rtol=1e-14)


def _two_stream_column(fn):
    """Oil enters at the top perforation and water at the bottom one, the
    middle is idle: the first segment carries both streams, (800 + 1020) / 2
    = 910 kg/m3, the two below it carry the water alone (1020 kg/m3)."""
    g = 9.80665
    dz = np.array([2.0, 5.0, 9.0])
    rho = np.array([[1000.0, 800.0], [1010.0, 790.0], [1020.0, 780.0]])
    flux = np.array([[0.0, -1.0], [0.0, 0.0], [-1.0, 0.0]])
    cdp = fn(dz, np.ones(3), rho, np.full((3, 2), 0.5), [0, 1], -1, flux, g)
    want = g * np.array([910.0 * 2, 910.0 * 2 + 1020.0 * 3, 910.0 * 2 + 1020.0 * 7])
    np.testing.assert_allclose(cdp, want, rtol=1e-14)


@pytest.mark.parametrize("impl", [0, 1])
def test_column_two_streams_hand_computed(impl):
    _two_stream_column(_column_impls()[impl])


@pytest.mark.parametrize("impl", [0, 1])
def test_column_injector_and_producer_without_history(impl):
    fn = _column_impls()[impl]
    g = 9.80665
    dz = np.array([3.0])
    rho = np.array([[1000.0, 800.0]])
    mob = np.array([[0.2, 0.6]])
    cdp = fn(dz, np.array([2.0]), rho, mob, [0.25, 0.75], 1, None, g)
    assert cdp[0] == pytest.approx(g * (0.25 * 1000 + 0.75 * 800) * 3.0)
    cdp = fn(dz, np.array([2.0]), rho, mob, [0.0, 1.0], -1, None, g)
    assert cdp[0] == pytest.approx(g * (0.2 * 1000 + 0.6 * 800) / 0.8 * 3.0)


def test_column_product_equals_oracle_on_random_columns():
    from pymrst_amd.generic import column_drops
    rng = np.random.default_rng(4)
    for trial in range(50):
        n = int(rng.integers(1, 7))
        dz = np.sort(rng.uniform(0.0, 30.0, n))
        wi = rng.uniform(0.5, 2.0, n)
        rho = rng.uniform(700.0, 1100.0, (n, 2))
        mob = rng.uniform(0.0, 1.0, (n, 2))
        flux = rng.normal(0.0, 1.0, (n, 2)) * (rng.random((n, 1)) > 0.3)
        sign = int(rng.choice([-1, 1]))
        a = column_pressure_drop(dz, wi, rho, mob, [0.3, 0.7], sign, flux, 9.81)
        b = column_drops(dz, wi, rho, mob, [0.3, 0.7], sign, flux, 9.81)
        np.testing.assert_allclose(b, a, rtol=1e-13, atol=1e-9)


# -- rhoScale, balance, Jacobian ---------------------------------------------------
def _gravity_case(mod=None):
    """A 4x4x5 column with gravity, a rate injector perforated in three
    cells and a bhp producer in four (it takes water in at its top)."""
    M = GenericOWModel if mod is None else mod.GenericOWModel
    g = og.cart_grid([4, 4, 5], [80.0, 80.0, 25.0])
    rng = np.random.default_rng(2)
    rock = og.make_rock(g, np.exp(rng.normal(np.log(200 * MD), 0.5, (80, 1))), np.full(80, 0.2))
    grav = np.array([0, 0, og.NORM_GRAVITY])
    col = lambda i, j, ks: np.array([i + 4 * (j + 4 * k) for k in ks])
    wells = og.add_well(None, g, rock, col(0, 0, [2, 3, 4]), "rate", 2e-4, compi=np.array([1.0, 0.0]), sign=1,
                        gravity=grav)
    wells = og.add_well(wells, g, rock, col(3, 3, [0, 1, 2, 3]), "bhp", 195 * BARSA, compi=np.array([0.0, 1.0]),
                        sign=-1, gravity=grav, lims={"lrat": -1e-4})
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    m = M(g, rock, _fluid(c_r=1e-10 / BARSA), wells, N, T, gravity=og.NORM_GRAVITY)
    st = m.initial_state(200 * BARSA, 0.3)
    st.pressure = st.pressure + 1000.0 * og.NORM_GRAVITY * g.cell_centroids[:, 2]
    return m, st


def test_rho_scale_is_taken_over_the_grid_at_the_step_start():
    m, s0 = _gravity_case()
    dp = 30 * BARSA
    s0.pressure = np.full(m.nc, 200 * BARSA + dp)
    st = s0.copy()
    st.pressure = st.pressure + 5 * BARSA
    cdp = m.compute_cdp_per_well(s0)
    _, _, eqs = m.residual_and_jacobian(s0, st, DAY, cdp)
    raw = GenericOWModel.__new__(GenericOWModel)
    raw.__dict__.update(m.__dict__)
    raw._rho_scale = np.ones(2)
    _, _, eqs1 = raw.residual_and_jacobian(s0, st, DAY, cdp)
    f = m.fluid
    for i, row in ((0, 2), (1, 3)):
        scale = value(eqs[row]) / value(eqs1[row])
        np.testing.assert_allclose(scale, np.exp(-f.c[i] * dp), rtol=1e-12)


def test_a_converged_step_balances_the_wells():
    m, s0 = _gravity_case()
    dt = 5 * DAY
    st, ok, its, _ = StabilisedNewton().step(m, s0, dt)
    assert ok
    f = m.fluid
    for i, s_new, s_old, q in ((0, st.sW, s0.sW, st.qWs), (1, 1 - st.sW, 1 - s0.sW, st.qOs)):
        inplace = lambda p, s: np.sum(m.pv * value(f.pv_mult(p)) * value(f.b(i, p)) * s)
        change = inplace(st.pressure, s_new) - inplace(s0.pressure, s_old)
        assert change == pytest.approx(dt * q.sum(), rel=1e-6)


def test_facility_jacobian_equals_central_differences():
    m, s0 = _gravity_case()
    st = s0.copy()
    rng = np.random.default_rng(8)
    st.pressure = st.pressure + rng.normal(0, 1e5, m.nc)
    st.sW = np.clip(st.sW + rng.normal(0, 0.05, m.nc), 0.01, 0.99)
    st.bhp = np.array([230 * BARSA, 190 * BARSA])     # the producer's top perforation injects
    st.qWs = np.array([2e-4, -5e-5])
    st.qOs = np.array([0.0, -1e-4])
    dt = 3 * DAY
    cdp = m.compute_cdp_per_well(s0)
    r0, J, _ = m.residual_and_jacobian(s0, st, dt, cdp)
    x0 = np.concatenate([st.pressure, st.sW, st.qWs, st.qOs, st.bhp])
    nc, nw = m.nc, 2
    Jd = J.toarray()
    cols = list(rng.choice(2 * nc, 12, replace=False)) + list(range(2 * nc, 2 * nc + 3 * nw))
    for c in cols:
        h = 1e-7 * max(abs(x0[c]), 1.0)
        def res(x):
            s = s0.copy()
            s.pressure, s.sW = x[:nc].copy(), x[nc:2 * nc].copy()
            s.qWs, s.qOs, s.bhp = x[2 * nc:2 * nc + nw].copy(), x[2 * nc + nw:2 * nc + 2 * nw].copy(), \
                x[2 * nc + 2 * nw:].copy()
            return m.residual_and_jacobian(s0, s, dt, cdp)[0]
        xp, xm = x0.copy(), x0.copy()
        xp[c] += h
        xm[c] -= h
        fd = (res(xp) - res(xm)) / (2 * h)
        scale = np.maximum(np.abs(Jd[:, c]), np.abs(fd)).max()
        assert np.abs(Jd[:, c] - fd).max() <= 1e-5 * max(scale, 1e-30), c


# -- planted mistakes ----------------------------------------------------------
def _check_two_streams():
    import oracle.generic as mod
    _two_stream_column(mod.column_pressure_drop)


def _check_rules():
    import oracle.generic as mod
    _switch_rules(mod.violated_limit)


def _check_wrong_way():
    import oracle.generic as mod
    w = _w(-1, None, "bhp", 1.5e7)
    assert float(np.ravel(value(mod.control_equation(w, _qs(2.0, 3.0), np.array([1e7]))))[0]) == 5.0


def _check_zero_target():
    import oracle.generic as mod
    w = _w(-1, None, "orat", 0.0)
    assert float(np.ravel(value(mod.control_equation(w, _qs(-2.0, -3.0), np.array([1e7]))))[0]) == -5.0


def _check_scaled():
    m, s0 = _gravity_case()
    s0.pressure = np.full(m.nc, 230 * BARSA)
    m.compute_cdp_per_well(s0)
    assert abs(m._rho_scale[1] - np.exp(-m.fluid.c[1] * 30 * BARSA)) < 1e-12


def test_planted_mistakes_are_caught():
    import oracle.generic as mod
    plants = [
        (mod, "column_pressure_drop",
         ("m_down = np.abs(np.cumsum(q_mass[::-1])[::-1])", "m_down = np.abs(np.cumsum(q_mass))"), _check_two_streams),
        (mod, "violated_limit", ('("orat", qo < lims.get("orat", -np.inf)),', ""), _check_rules),
        (mod, "violated_limit", ('order = [("bhp", bhp < lims.get("bhp", -np.inf)), ',
                                 'order = [("bhp", bhp > lims.get("bhp", -np.inf)), '), _check_rules),
        (mod, "control_equation", ("and np.sign(qv) != w.sign):", "and False):"), _check_wrong_way),
        (mod, "control_equation", ("    if float(w.val) == 0.0:\n        return qt\n", ""), _check_zero_target),
        (mod.GenericOWModel, "compute_cdp_per_well",
         ("np.array([f.rhoS[i] / np.mean(f.rhoS[i] * b0[i]) for i in range(2)])", "np.ones(2)"), _check_scaled),
    ]
    for obj, name, edit, check in plants:
        check()
        with planted(obj, name, edit):
            assert fails(check), f"planted mistake not caught: {edit[0]!r}"
