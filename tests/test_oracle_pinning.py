"""Known answers that pin the oracle where no golden run reaches: oil-water
gravity and the well-bore head (cdp), rock compressibility, the single-phase
compressible Newton of MRST's flowSolverTutorialAD, the black-oil flash and
well balances, the CPR solver of the CPU baseline, and the report-step cut
rule.  Every check runs once on the oracle as it is (it must pass) and once
with a plausible mistake planted in the oracle (it must fail), so no check
is vacuous.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle.bo as obo
import oracle.grid as og
import oracle.nonlinear as onl
import oracle.ow as oow
from oracle.adi import init_variables_adi, value
from oracle.cpr import CPRSolverCPU

from cases import golden, paired_inputs, build_oracle
from mutation import fails, planted

G = og.NORM_GRAVITY
BAR = og.BARSA
MU = np.array([1.0, 5.0]) * og.CENTI * og.POISE


def _column(nz, height, c, *, gravity, wells, c_r=None, rho=(1000.0, 700.0), **kw):
    """A 1 x 1 x nz column of 10 x 10 m cells, 100 mD, porosity 0.25, with
    wells given as (cells, type, val, compi, sign)."""
    g = og.cart_grid([1, 1, nz], [10.0, 10.0, height])
    rock = og.make_rock(g, 100 * og.MILLI * og.DARCY, 0.25)
    gv = np.array([0.0, 0.0, G])
    ws = None
    for cells, kind, val, compi, sign in wells:
        ws = og.add_well(ws, g, rock, np.array(cells), kind, val, compi=np.array(compi), sign=sign, gravity=gv)
    fl = oow.OWFluid(mu=MU, rhoS=np.array(rho), n=np.array([2.0, 2.0]), c=np.asarray(c, float),
                     p_ref=200 * BAR, c_r=c_r)
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    return oow.OWModel(g, rock, fl, ws, N, T, gravity=gravity, **kw)


# -- oil-water gravity (two_phase_oil_water_model.py:738-891) ------------------
def check_hydrostatic_water_column_has_zero_flux():
    """A water column at hydrostatic equilibrium, p = p_ref + rho_w g z with
    sW = 1 and no compressibility, carries no flux: the gravity term of the
    phase potential cancels the pressure gradient (:738-773).  A zero-WI well
    keeps the model callable without touching the column."""
    m = _column(5, 50.0, [0.0, 0.0], gravity=G, wells=[([0], "rate", 0.0, [1.0, 0.0], 1)])
    m.wells[0].WI = np.array([0.0])
    z = m.g.cell_centroids[:, 2]
    p = 200 * BAR + 1000.0 * G * z
    sw = np.ones(5)
    st0 = oow.OWState(p.copy(), sw.copy(), np.zeros(1), np.zeros(1), np.array([p[0]]))
    water, oil, *_ = m.assemble(st0, *init_variables_adi(p, sw, [0.0], [0.0], [p[0]]), 1.0)
    assert np.abs(value(water)).max() <= 1e-8, value(water)
    assert np.abs(value(oil)).max() <= 1e-8


def _bottom_sw_after_injection(gravity):
    m = _column(6, 60.0, [1e-6 / BAR, 1e-6 / BAR], gravity=gravity,
                wells=[([0], "rate", 2e-6, [1.0, 0.0], 1), ([5], "bhp", 190 * BAR, [0.5, 0.5], -1)])
    st = m.initial_state(200 * BAR, 0.2)
    for _ in range(8):
        st, ok, _ = m.step(st, 5 * og.DAY)
        assert ok
    return st.sW[-1]


def check_gravity_speeds_denser_water_downward():
    """Water (1000 kg/m3) injected at the top of a column over oil
    (700 kg/m3), producer at the bottom: after 8 x 5 days the bottom cell
    holds more water with gravity than without (:775-810)."""
    sw_g, sw_0 = _bottom_sw_after_injection(G), _bottom_sw_after_injection(0.0)
    assert sw_g > sw_0 + 0.01, (sw_g, sw_0)


def check_segment_mix_fractions():
    """Perforation rates (1, 2, 0) water and (0.5, 0, 0) oil, top to bottom:
    the segments carry (3, 2, 0) water and (0.5, 0, 0) oil; the dead bottom
    segment takes compi (0.3, 0.7) (:812-848)."""
    mw, mo = oow.segment_mix_fractions(np.array([0.3, 0.7]), np.array([1.0, 2.0, 0.0]), np.array([0.5, 0.0, 0.0]))
    assert np.allclose(mw, [3.0 / 3.5, 1.0, 0.3], rtol=1e-15, atol=0)
    assert np.allclose(mo, [0.5 / 3.5, 0.0, 0.7], rtol=1e-15, atol=0)


def check_topology_cdp():
    """Two perforations 5 m and 15 m below the reference depth with mixtures
    0.8/0.2 and 0.3/0.7 (b = 1, rho 1000/700): densities 940 and 790, so
    cdp = (g 940 5, g 940 5 + g 790 10) = (46,091.255, 123,563.79) Pa; zero
    without gravity (:850-891).  With compressed phases a pure phase weighs
    its reservoir density rhoS b: pure water at b_w = 1.02 over pure oil at
    b_o = 1.25 gives 1020 and 875 kg/m3.  The well's dZ comes from add_well."""
    m = _column(2, 20.0, [0.0, 0.0], gravity=G, wells=[([0, 1], "bhp", 190 * BAR, [0.5, 0.5], -1)])
    w = m.wells[0]
    assert np.array_equal(w.dZ, [5.0, 15.0])
    mw, mo, one = np.array([0.8, 0.3]), np.array([0.2, 0.7]), np.ones(2)
    cdp = oow.topology_cdp(w.dZ, mw, mo, one, one, m.fluid.rhoS, G)
    assert np.allclose(cdp, [9.80665 * 4700.0, 9.80665 * 4700.0 + 9.80665 * 7900.0], rtol=1e-15, atol=0)
    assert np.array_equal(oow.topology_cdp(w.dZ, mw, mo, one, one, m.fluid.rhoS, 0.0), [0.0, 0.0])
    cdp = oow.topology_cdp(w.dZ, np.array([1.0, 0.0]), np.array([0.0, 1.0]), np.full(2, 1.02), np.full(2, 1.25),
                           m.fluid.rhoS, G)
    assert np.allclose(cdp, [9.80665 * 5100.0, 9.80665 * 5100.0 + 9.80665 * 8750.0], rtol=1e-14, atol=0)


def check_cdp_is_frozen_for_the_step():
    """step computes the well heads once from the start-of-step state and
    keeps them through every Newton iteration (:893-930)."""
    m = _column(6, 60.0, [1e-6 / BAR, 1e-6 / BAR], gravity=G,
                wells=[([0], "rate", 2e-6, [1.0, 0.0], 1), ([5], "bhp", 190 * BAR, [0.5, 0.5], -1)])
    calls = []
    orig = m.compute_cdp_per_well
    m.compute_cdp_per_well = lambda st0: calls.append(1) or orig(st0)
    _, ok, its = m.step(m.initial_state(200 * BAR, 0.2), 5 * og.DAY)
    assert ok and its >= 2 and len(calls) == 1, (its, len(calls))


def check_rock_compressibility_closed_form():
    """Two water-filled cells (pv 250 m3 each, c_w = 0 so b_w = 1) with rock
    compressibility c_r and an injector of q m3/s: since pvMultR(p) =
    1 + c_r (p - p_ref) (init_simple_adi_fluid.py:313-320), the stored water
    pv sum(1 + c_r (p_i - p_ref)) grows by q dt, so the mean pressure rises
    by exactly q dt / (2 pv c_r) = 8.64 / (500e-8) Pa = 17.28 bar."""
    c_r, q, dt = 1e-8, 1e-4, og.DAY
    m = _column(2, 20.0, [0.0, 1e-9], gravity=0.0, wells=[([0], "rate", q, [1.0, 0.0], 1)], c_r=c_r,
                tol_cnv=1e-12, tol_mb=1e-15, tol_well=1e-13)
    st0 = m.initial_state(200 * BAR, 1.0)
    st, ok, _ = m.step(st0, dt)
    assert ok
    dp_mean = np.mean(st.pressure - st0.pressure)
    assert abs(dp_mean - q * dt / (500.0 * c_r)) <= 1e-9 * dp_mean, dp_mean
    assert abs(dp_mean - 17.28 * BAR) <= 1e-9 * dp_mean


def check_compressible_tutorial_matches_mrst():
    """MRST's flowSolverTutorialAD (ref_p4_ad.mat): hydrostatic start, well
    index and depth of a y-direction well, 53 pressure fields, bhp, rates
    and per-step Newton counts (examples/validation/validate_phase4.py:35-45
    tolerances)."""
    from oracle import compressible
    out = compressible.run_simulation()
    ref = golden("p4_ad.npz")
    scale = 250 * BAR
    assert np.abs(out["p_init"] - ref["p_init"]).max() <= 1e-9 * scale
    assert np.abs(out["WI"] - ref["WI"]).max() <= 1e-12 * ref["WI"].max()
    assert np.abs(out["dZ"] - ref["dZ"]).max() <= 1e-12 * 50
    assert np.abs(out["p_hist"] - ref["p_hist"]).max() <= 1e-8 * scale
    assert np.abs(out["bhp_hist"] - ref["bhp_hist"]).max() <= 1e-8 * scale
    assert np.abs(out["qs_hist"] - ref["qs_hist"]).max() <= 1e-8 * np.abs(ref["qs_hist"]).max()
    assert np.array_equal(out["nits"], ref["nits"].astype(int))


OW_CHECKS = {
    "hydrostatic": (check_hydrostatic_water_column_has_zero_flux,
                    (oow.OWModel, "assemble", ("pot_w = dp - ops.avg(rho_w)", "pot_w = dp + ops.avg(rho_w)"))),
    "gravity_front": (check_gravity_speeds_denser_water_downward,
                      (oow.OWModel, "assemble", ("pot_w = dp - ops.avg(rho_w)", "pot_w = dp + ops.avg(rho_w)"),
                       ("pot_o = dp - ops.avg(rho_o)", "pot_o = dp + ops.avg(rho_o)"))),
    "segment_mix": (check_segment_mix_fractions,
                    (oow, "segment_mix_fractions", ("np.cumsum(np.asarray(cq_w, float)[::-1])[::-1]",
                                                    "np.cumsum(np.asarray(cq_w, float))"))),
    "topology_cdp": (check_topology_cdp,
                     (oow, "topology_cdp", ("(mix_w * rhoS[0] + mix_o * rhoS[1]) / vr",
                                            "(mix_w * rhoS[0] + mix_o * rhoS[1])"))),
    "cdp_frozen": (check_cdp_is_frozen_for_the_step,
                   (oow.OWModel, "step", ("ok, _ = self.converged(eqs, dt, st)",
                                          "cdp = self.compute_cdp_per_well(st0); ok, _ = self.converged(eqs, dt, st)"))),
    "rock_compressibility": (check_rock_compressibility_closed_form,
                             (oow.OWFluid, "pv_mult", ("1.0 + self.c_r * (p - self.p_ref)",
                                                       "1.0 + self.c_r * (p - self.p_ref) ** 1.0 * 0.5"))),
    "compressible_tutorial": (check_compressible_tutorial_matches_mrst,
                              (oow.Ops, "avg", ("return ad_matmul(self.M, x)",
                                                "return ad_matmul(self.M, x) * 0.999"))),
}


@pytest.mark.parametrize("name", list(OW_CHECKS))
def test_oil_water_known_answer(name):
    OW_CHECKS[name][0]()


@pytest.mark.parametrize("name", list(OW_CHECKS))
def test_oil_water_known_answer_sees_a_planted_mistake(name):
    check, (obj, attr, *edits) = OW_CHECKS[name]
    with planted(obj, attr, *edits):
        assert fails(check), f"{name}: the planted mistake went unnoticed"


# -- black-oil flash (three_phase_black_oil_model.py:472-509) -------------------
# The known answers of tests/ad_blackoil/test_compute_flash_black_oil.py,
# mapped onto the model's flash: its previous-status argument is the
# reference test's state0 status (0 there means "not undersaturated", i.e.
# status 3 here), oil is always present (no vaporised oil) and dissolved gas
# is always on, so the ~disgas cases do not apply.
class _FlatRsFluid:
    """Flat saturation curve Rs_sat = 50, so the numbers come from the flash."""

    def rs_sat(self, p):
        return np.full(np.size(p), 50.0)


def _flash(s, rs, prev_s, prev_rs=10.0, prev_status=3, p=2.0e7):
    m = object.__new__(obo.BOModel)
    m.fluid = _FlatRsFluid()
    s = np.atleast_2d(np.asarray(s, float))
    n = s.shape[0]
    prev_s = np.broadcast_to(np.atleast_2d(np.asarray(prev_s, float)), (n, 3)).copy()
    z = np.zeros(1)
    st = obo.BOState(np.full(n, p), s.copy(), np.broadcast_to(np.asarray(rs, float), (n,)).copy(),
                     np.full(n, 3, np.int64), z, z, z, z)
    prev = obo.BOState(np.full(n, p), prev_s, np.broadcast_to(np.asarray(prev_rs, float), (n,)).copy(),
                       np.full(n, prev_status, np.int64), z, z, z, z)
    m.flash(st, prev, prev.status == 1)
    return st


def check_flash_clips_negative_gas():
    """s = (0.2, 0.9, -0.1): gas clipped, the others divided by 1.1."""
    st = _flash([0.2, 0.9, -0.1], 10.0, [0.2, 0.8, 0.0])
    assert np.allclose(st.s, [[0.2 / 1.1, 0.9 / 1.1, 0.0]], rtol=1e-12, atol=0) and st.s[0, 2] == 0.0
    for sats in ([-0.1, 0.6, 0.5], [0.3, -0.1, 0.8], [0.3, 0.8, -0.1]):
        row = _flash(sats, 10.0, [0.3, 0.4, 0.3]).s[0]
        assert np.all(row >= 0.0) and abs(row.sum() - 1.0) <= 1e-12, (sats, row)
    for sats in ([0.2, 0.9, -0.1], [-0.1, 1.2, -0.1], [0.3, -0.2, 0.9]):
        clamped = np.maximum(np.array(sats), 0.0)
        assert np.allclose(_flash(sats, 10.0, [0.3, 0.4, 0.3]).s[0], clamped / clamped.sum(), rtol=1e-12, atol=0)
    st = _flash([0.2, 0.5, 0.1], 10.0, [0.2, 0.5, 0.1])
    assert np.allclose(st.s, [[0.25, 0.625, 0.125]], rtol=1e-15, atol=0)


def check_flash_sets_rs_on_the_curve_with_gas_and_caps_it_without():
    """Free gas present -> rs set to Rs_sat (10 -> 50); no gas and rs above
    the curve -> capped (90 -> 50); no gas below the curve -> kept (10)."""
    st = _flash([[0.1, 0.6, 0.3], [0.1, 0.9, 0.0], [0.1, 0.9, 0.0]], [10.0, 90.0, 10.0],
                [[0.1, 0.6, 0.3], [0.1, 0.9, 0.0], [0.1, 0.9, 0.0]])
    assert np.array_equal(st.rs, [50.0, 50.0, 10.0]), st.rs
    assert np.array_equal(st.status, [3, 1, 1])


def check_flash_keeps_rs_non_negative():
    """An undersaturated cell without free gas at rs = -5 ends at rs = 0,
    status 1."""
    st = _flash([0.2, 0.8, 0.0], -5.0, [0.2, 0.8, 0.0], prev_rs=-5.0, prev_status=1)
    assert st.status[0] == 1 and st.rs[0] == 0.0


def check_flash_status_packs_gas_presence():
    """status = 1 (oil present) + 2 (gas present): free gas gives 3, none 1."""
    assert _flash([0.1, 0.6, 0.3], 10.0, [0.1, 0.6, 0.3]).status[0] == 3
    assert _flash([0.1, 0.9, 0.0], 10.0, [0.1, 0.9, 0.0]).status[0] == 1


def check_flash_water_only_cell_is_saturated():
    """sW = 1 forces gas presence (status 3) whatever the previous status."""
    for prev_status in (1, 3):
        assert _flash([1.0, 0.0, 0.0], 10.0, [1.0, 0.0, 0.0], prev_status=prev_status).status[0] == 3


def check_flash_phase_appearance_rules():
    """Gas appears in an undersaturated cell only when rs crosses the curve
    from a start on it: rs 60 from rs0 = 50 = Rs_sat -> saturated (status 3,
    rs 50, sG 0); rs 60 from rs0 = 40 -> stays undersaturated with rs capped
    at 50.  A saturated cell whose sG goes negative keeps its gas (status 3,
    sG clipped to 0, rs on the curve)."""
    st = _flash([0.2, 0.8, 0.0], 60.0, [0.2, 0.8, 0.0], prev_rs=50.0, prev_status=1)
    assert st.status[0] == 3 and st.rs[0] == 50.0 and st.s[0, 2] == 0.0
    st = _flash([0.2, 0.8, 0.0], 60.0, [0.2, 0.8, 0.0], prev_rs=40.0, prev_status=1)
    assert st.status[0] == 1 and st.rs[0] == 50.0
    st = _flash([0.2, 0.85, -0.05], 20.0, [0.2, 0.7, 0.1], prev_rs=50.0, prev_status=3)
    assert st.status[0] == 3 and st.rs[0] == 50.0 and st.s[0, 2] == 0.0
    assert np.allclose(st.s[0], [0.2 / 1.05, 0.85 / 1.05, 0.0], rtol=1e-14, atol=0)


FLASH_CHECKS = {
    "clipping": (check_flash_clips_negative_gas, ("st.s = s / s.sum(axis=1, keepdims=True)", "st.s = s")),
    "rs_set_or_capped": (check_flash_sets_rs_on_the_curve_with_gas_and_caps_it_without,
                         ("rs[gas] = rs_sat[gas]", "rs[gas] = np.minimum(rs_sat[gas], rs[gas])")),
    "rs_non_negative": (check_flash_keeps_rs_non_negative, ("st.rs = np.maximum(rs, 0.0)", "st.rs = rs")),
    "status": (check_flash_status_packs_gas_presence, ("st.status = 1 + 2 * gas", "st.status = 1 + gas")),
    "water_only": (check_flash_water_only_cell_is_saturated, ("gas |= wat_only", "gas = gas")),
    "appearance": (check_flash_phase_appearance_rules, ("gas |= ix2", "gas = gas")),
}


@pytest.mark.parametrize("name", list(FLASH_CHECKS))
def test_black_oil_flash_known_answer(name):
    FLASH_CHECKS[name][0]()


@pytest.mark.parametrize("name", list(FLASH_CHECKS))
def test_black_oil_flash_known_answer_sees_a_planted_mistake(name):
    check, edit = FLASH_CHECKS[name]
    with planted(obo.BOModel, "flash", edit):
        assert fails(check), f"{name}: the planted mistake went unnoticed"


# -- black-oil wells (three_phase_black_oil_model.py:383-436, 1040-1060) --------
def _small_bo():
    from oracle.cases import bo_model_from_case
    from pymrst_amd.workloads import blackoil_field
    c = blackoil_field(dims=(6, 6, 5), cap_layers=2, steps=2, step_days=5.0, gas_rate=2e4 / og.DAY)
    m, st = bo_model_from_case(c)
    m.tol_cnv, m.tol_mb, m.tol_well = 1e-9, 1e-12, 1e-10
    return c, m, st


def _in_place(m, st):
    f, p = m.fluid, st.pressure
    pv = m.pv_ref * f.pv_mult(p)
    bo = f.b_o(p, st.rs, st.status == 3)
    return np.array([np.sum(pv * f.b_w(p) * st.s[:, 0]), np.sum(pv * bo * st.s[:, 1]),
                     np.sum(pv * (f.b_g(p) * st.s[:, 2] + st.rs * bo * st.s[:, 1]))])


def check_black_oil_surface_volumes_balance_the_wells():
    """A gas injector (2e4 Sm3/day, pure gas) in a gas cap and four bhp
    producers, two converged 5-day steps: the surface water, oil and gas
    (free + dissolved) in place change by exactly dt times the wells' net
    surface rates (the reference's gas-balance test, :1040-1060, for all
    three phases), and the injector meets its target with pure gas."""
    c, m, st = _small_bo()
    v0, net = _in_place(m, st), np.zeros(3)
    for dt in c.dts:
        st, ok, _ = m.step(st, dt)
        assert ok
        assert st.qGs[0] == pytest.approx(2e4 / og.DAY, rel=1e-9) and st.qWs[0] == 0.0 and st.qOs[0] == 0.0
        net += dt * np.array([st.qWs.sum(), st.qOs.sum(), st.qGs.sum()])
    change = _in_place(m, st) - v0
    assert np.all(np.abs(change - net) <= 1e-9 * np.abs(net)), (change, net)


def _perforation_rates(m, k, p, bhp, rs, sg):
    """Connection surface rates (water, oil, gas) of well k at a state of
    uniform pressure p, rs per cell, free gas saturation sg per cell."""
    f, nc = m.fluid, m.nc
    p = np.full(nc, p)
    sw = np.full(nc, 0.2)
    so = 1.0 - sw - sg
    sat = sg > 0
    b = (value(f.b_w(p)), value(f.b_o(p, rs, sat)), value(f.b_g(p)))
    mob = (value(f.kr_w(sw)) / value(f.mu_w(p)), value(f.kr_o(sw, so, sg)) / value(f.mu_o(p, rs, sat)),
           value(f.kr_g(sg)) / value(f.mu_g(p)))
    z = np.zeros(1)
    cq, _, _ = m._well(k, p, mob, b, rs, (z, z, z), np.array([bhp]))
    return [value(x) for x in cq], b, mob


def check_black_oil_well_rates():
    """Producer in undersaturated oil (rs 40 and 60 Sm3/Sm3 at its two
    perforations, no free gas): each connection's surface gas is rs times its
    surface oil, the solution gas the oil carries.  With free gas at one
    perforation that connection's gas exceeds rs times its oil.  A pure gas
    injector: no water or oil, and surface gas = b_g x the reservoir inflow
    WI (sum of mobilities) (bhp - p)."""
    _, m, _ = _small_bo()
    prod = m.wells[1].cells
    rs = np.full(m.nc, 40.0)
    rs[prod[1]] = 60.0
    sg = np.zeros(m.nc)
    cq, _, _ = _perforation_rates(m, 1, 200 * BAR, 150 * BAR, rs, sg)
    assert np.all(cq[1] < 0)
    assert np.allclose(cq[2], np.array([40.0, 60.0]) * cq[1], rtol=1e-14, atol=0)
    sg[prod[0]] = 0.1
    cq, _, _ = _perforation_rates(m, 1, 200 * BAR, 150 * BAR, rs, sg)
    assert cq[2][0] < 40.0 * cq[1][0] * (1 + 1e-3) and cq[2][1] == pytest.approx(60.0 * cq[1][1], rel=1e-14)
    inj = m.wells[0]
    cq, b, mob = _perforation_rates(m, 0, 200 * BAR, 210 * BAR, np.full(m.nc, 40.0), np.full(m.nc, 0.25))
    assert np.all(cq[0] == 0.0) and np.all(cq[1] == 0.0)
    expect = b[2][inj.cells] * inj.WI * (mob[0] + mob[1] + mob[2])[inj.cells] * (10 * BAR)
    assert np.allclose(cq[2], expect, rtol=1e-13, atol=0)


BO_WELL_CHECKS = {
    "balance": (check_black_oil_surface_volumes_balance_the_wells,
                (obo.BOModel, "assemble", ("sub = -cq[ph]", "sub = -cq[ph] * (1.0 - 1e-3 * (ph == 2))"))),
    "rates_gor": (check_black_oil_well_rates,
                  (obo.BOModel, "_well", ("cq_p[2] + rs_c * (b_c[1] * cq_p[1])", "cq_p[2]"))),
    "rates_injection": (check_black_oil_well_rates,
                        (obo.BOModel, "_well", ("mix[ph] * cqt_is for ph", "mix[ph] * cqt_i for ph"))),
}


@pytest.mark.parametrize("name", list(BO_WELL_CHECKS))
def test_black_oil_well_known_answer(name):
    BO_WELL_CHECKS[name][0]()


@pytest.mark.parametrize("name", list(BO_WELL_CHECKS))
def test_black_oil_well_known_answer_sees_a_planted_mistake(name):
    check, (obj, attr, *edits) = BO_WELL_CHECKS[name]
    with planted(obj, attr, *edits):
        assert fails(check), f"{name}: the planted mistake went unnoticed"


# -- the CPR solver of the CPU baseline (linear_solvers.py:110-545, 595-807) ----
def _cpr_case(nx, ny, nz, p=200 * BAR):
    """The reference's CPR test case (linear_solvers.py:601-628): a
    500 x 500 x 50 m box, k = 10 + 500 (1 + sin(3x/90) cos(3y/70)) mD,
    a water injector of 5e-3 m3/s in the first cell and a 150 bar producer
    in the last; the Jacobian of a 30-day step at the initial state."""
    g = og.cart_grid([nx, ny, nz], [500.0, 500.0, 50.0])
    c = g.cell_centroids
    k = (10 + 500 * (1 + np.sin(3 * c[:, 0] / 90.0) * np.cos(3 * c[:, 1] / 70.0))) * og.MILLI * og.DARCY
    rock = og.make_rock(g, k, 0.25)
    fl = oow.OWFluid(mu=MU, rhoS=np.array([1000.0, 800.0]), n=np.array([2.0, 2.0]),
                     c=np.array([1e-6 / BAR, 1e-4 / BAR]), p_ref=200 * BAR)
    nc = g.num_cells
    w = og.add_well(None, g, rock, [0], "rate", 5e-3, compi=np.array([1.0, 0.0]), sign=1)
    w = og.add_well(w, g, rock, [nc - 1], "bhp", 150 * BAR, compi=np.array([0.0, 1.0]), sign=-1)
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    m = oow.OWModel(g, rock, fl, w, N, T)
    st = oow.OWState(np.full(nc, p), np.full(nc, 0.2), np.array([5e-3, 0.0]), np.array([0.0, -5e-3]),
                     np.array([p, 150 * BAR]))
    r, J, _ = m.residual_and_jacobian(st, st.copy(), 30 * og.DAY)
    return m, st, J.tocsr(), -r


def check_cpr_reproduces_the_direct_solution():
    """CPR = direct to 1e-7 (linear_solvers.py:639-649), and in at most 12
    GMRES iterations on the 10 x 10 x 3 case (it takes 9; a preconditioner
    missing its second-stage residual correction takes 23)."""
    m, st, J, rhs = _cpr_case(10, 10, 3)
    exact = spla.spsolve(J.tocsc(), rhs)
    s = CPRSolverCPU(rtol=1e-10)
    got = s.solve(J, rhs, m, {"pressure": st.pressure})
    assert np.abs(got - exact).max() <= 1e-7 * np.abs(exact).max()
    assert s.last_iterations <= 12, s.last_iterations


def check_pressure_system_is_an_m_matrix_and_decoupled():
    """W A P is an M-matrix: positive diagonal, diagonally dominant, no
    positive off-diagonal entries (linear_solvers.py:651-680).  And the
    quasi-IMPES weights 1/b at the mean pressure decouple it: at a uniform
    pressure (250 bar, away from p_ref) without gravity the weighted
    equations do not depend on the saturations, W A[:, sW] = 0, except in
    the perforated cells."""
    m, st, J, _ = _cpr_case(20, 20, 5, p=250 * BAR)
    nc, n = m.nc, J.shape[0]
    w = CPRSolverCPU()._weighting(m, {"pressure": st.pressure}, n)
    prolong = sp.csr_array((np.ones(nc), (np.arange(nc), np.arange(nc))), shape=(n, nc))
    a_p = (w @ J @ prolong).tocsr()
    d = a_p.diagonal()
    off = np.abs(a_p).sum(axis=1) - np.abs(d)
    offd = (a_p - sp.diags_array(d)).tocoo()
    assert (d > 0).all() and (np.abs(d) >= off - 1e-30).all() and (offd.data <= 1e-14).all()
    ws = (w @ J).tocsr()[:, nc:2 * nc].toarray()
    perf = np.concatenate([x.cells for x in m.wells])
    keep = np.setdiff1d(np.arange(nc), perf)
    assert np.abs(ws[np.ix_(keep, keep)]).max() <= 1e-12 * np.abs(d).max()


def check_pressure_stage_earns_its_keep():
    """Ablation (linear_solvers.py:712-735): the same local stage without the
    pressure stage needs more iterations, or does not converge."""
    m, st, J, rhs = _cpr_case(25, 25, 8)
    cpr = CPRSolverCPU(rtol=1e-8)
    cpr.solve(J, rhs, m, {"pressure": st.pressure})
    local = CPRSolverCPU(rtol=1e-8, use_pressure_stage=False)
    try:
        local.solve(J, rhs, m, {"pressure": st.pressure})
        n_local = local.last_iterations
    except RuntimeError:
        n_local = None
    assert n_local is None or cpr.last_iterations < n_local, (cpr.last_iterations, n_local)


def check_cpr_newton_matches_the_reference_run():
    """The reference's paired waterflood with CPR (rtol 1e-10) instead of the
    direct solve: the same Newton counts per report step as the reference's
    run and the same states (README.md:769-772, linear_solvers.py:784-807:
    a different linear solver must not change the answer); first 8 steps."""
    inp = paired_inputs()
    m, s0 = build_oracle(inp, linear_solver=CPRSolverCPU(rtol=1e-10))
    states, its = oow.simulate_schedule(s0, m, inp["dts"][:8])
    ref = golden("paired_waterflood.npz")
    assert np.array_equal(its, ref["newton_iterations"][:8])
    p = np.vstack([s.pressure for s in states])
    assert np.abs(p - ref["pressure"][:8]).max() <= 1e-6 * np.abs(ref["pressure"]).max()
    assert np.abs(np.vstack([s.sW for s in states]) - ref["sw"][:8]).max() <= 1e-6


CPR_CHECKS = {
    "cpr_direct": (check_cpr_reproduces_the_direct_solution,
                   (CPRSolverCPU, "solve", ("return x + local(r - a @ x)", "return x + local(r)"))),
    "m_matrix": (check_pressure_system_is_an_m_matrix_and_decoupled,
                 (CPRSolverCPU, "_weighting", ("w = np.array([1.0 / float(", "w = np.array([1.0 * float("))),
    "pressure_stage": (check_pressure_stage_earns_its_keep,
                       (CPRSolverCPU, "solve", ("x = prolong @ solve_p(w @ r)", "x = 0.0 * (prolong @ solve_p(w @ r))"))),
    "cpr_newton": (check_cpr_newton_matches_the_reference_run,
                   (CPRSolverCPU, "solve", ("restart=self.restart, maxiter=self.maxiter,", "restart=5, maxiter=1,"),
                    ("if info != 0:", "if False:"))),
}


@pytest.mark.parametrize("name", list(CPR_CHECKS))
def test_cpr_known_answer(name):
    CPR_CHECKS[name][0]()


@pytest.mark.parametrize("name", list(CPR_CHECKS))
def test_cpr_known_answer_sees_a_planted_mistake(name):
    check, (obj, attr, *edits) = CPR_CHECKS[name]
    with planted(obj, attr, *edits):
        assert fails(check), f"{name}: the planted mistake went unnoticed"


# -- the report-step cut rule (nonlinear_solver.py:621-702, 955-977) ------------
class _ScriptedModel:
    """A model whose state is the simulated time: a ministep of dt from t
    converges in 3 iterations when dt <= limit(t), and fails after 25."""

    def __init__(self, limit):
        self.limit = limit

    def step(self, t, dt):
        if dt <= self.limit(t) * (1 + 1e-12):
            return t + dt, True, 3
        return t, False, 25


def check_cut_rule():
    """The whole step first, then 2^k equal sub-steps from the start state for
    k = 1..6; a cut step reports the sum of its sub-steps' iterations, a
    failed one the first attempt's iterations and 6 cuts; the schedule stops
    at the first report step that fails."""
    D = 8.0
    # converges only at dt <= D/4: D fails, 2 x D/2 fails at once, 4 x D/4 passes
    t, ok, its, cuts = onl.solve_timestep(_ScriptedModel(lambda t: D / 4), 0.0, D)
    assert (t, ok, its, cuts) == (D, True, 12, 2)
    # D/2 allowed before t = D/2, D/4 after: the first half-step of cut 1
    # succeeds and its second fails; cut 2 restarts from the start state
    t, ok, its, cuts = onl.solve_timestep(_ScriptedModel(lambda t: D / 2 if t < D / 2 else D / 4), 0.0, D)
    assert (t, ok, its, cuts) == (D, True, 12, 2)
    # never gets past D/2: every cut fails after its first half of sub-steps
    # (cut 6: 32 x 3 + 25 = 121 iterations), and the step reports the first
    # attempt's 25
    t, ok, its, cuts = onl.solve_timestep(_ScriptedModel(lambda t: D / 2 if t < D / 2 else 0.0), 0.0, D)
    assert (t, ok, its, cuts) == (D / 2, False, 25, 6)
    # the whole step converges: no cut
    assert onl.solve_timestep(_ScriptedModel(lambda t: D), 0.0, D) == (D, True, 3, 0)
    # the schedule stops after the report step that failed
    states, its, cuts = onl.simulate_schedule(_ScriptedModel(lambda t: D if t < 2 * D else 0.0), 0.0, [D, D, D, D])
    assert list(states) == [D, 2 * D, 2 * D] and list(its) == [3, 3, 25] and list(cuts) == [0, 0, 6]


CUT_MUTANTS = {
    "no_restart": (onl, "solve_timestep", ("st, ok, total = state0, True, 0", "ok, total = True, 0")),
    "failed_iterations": (onl, "solve_timestep", ("return st, False, its, max_cuts", "return st, False, total, max_cuts")),
    "no_stop": (onl, "simulate_schedule", ("        if not ok:\n            break\n", "")),
}


def test_cut_rule_known_answer():
    check_cut_rule()


@pytest.mark.parametrize("name", list(CUT_MUTANTS))
def test_cut_rule_sees_a_planted_mistake(name):
    obj, attr, *edits = CUT_MUTANTS[name]
    with planted(obj, attr, *edits):
        assert fails(check_cut_rule), f"{name}: the planted mistake went unnoticed"
