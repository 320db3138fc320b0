"""GPU parity of GenericBlackOilModelAD's three-phase live-oil form (a deck
fluid with dissolved gas, pymrst_amd.generic.GenericLiveOilModelAD) with
the FacilityModel wells, against the oracle restatement
(oracle.generic.GenericBOModel), both driven by the same Newton protocol
(NonLinearSolver / StabilisedNewton), through the C ABI
(pmx_bo_set_controls, pmx_bo_set_connection_drops, pmx_bo_set_update_rule,
pmx_bo_cell_densities, pmx_bo_perf_volume_rates).

Cases on a 4x4x5 column with gravity, a gas cap over undersaturated oil, a
gas injector perforated in the two top layers, a producer perforated in
four layers whose wellbore column is carried from step to step in its
perforation fluxes, and a shut producer: the producer on bhp switching to
its oil-rate limit; the producer on a gas-rate target; the injector asked
for more gas than its bhp limit buys, switching to bhp (gas moves into
undersaturated cells: flash status changes).  Bar: the same Newton
iterations and cuts per report step, the same control switches, the same
flash statuses, pressure and bhp within 1e-8 relative, saturations and rs
within 1e-8 of their scale, and the perforation fluxes handed on.
"""
import numpy as np
import pytest

import oracle.grid as og
from oracle.adi import value
from oracle.bo import BOFluid, synthetic_deck
from oracle.generic import GenericBOModel
from oracle.nonlinear import StabilisedNewton, simulate_stabilised

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BARSA, DAY = og.BARSA, og.DAY
MD = og.MILLI * og.DARCY


def _models(inj_rate=0.05, inj_lims=None, prod_type="bhp", prod_val=195 * BARSA, prod_lims=None):
    import pymrst_amd as pa
    from pymrst_amd.bo import init_deck_blackoil_fluid
    from pymrst_amd.generic import GenericBlackOilModelAD, GenericLiveOilModelAD
    from pymrst_amd.ow import CPRSolver
    dims, phys = [4, 4, 5], [80.0, 80.0, 25.0]
    rng = np.random.default_rng(2)
    perm = np.exp(rng.normal(np.log(200 * MD), 0.5, (80, 1)))
    poro = np.full(80, 0.2)
    grav = np.array([0, 0, og.NORM_GRAVITY])
    go, gd = og.cart_grid(dims, phys), pa.cart_grid(dims, phys)
    ro, rd = og.make_rock(go, perm, poro), pa.make_rock(gd, perm, poro)
    col = lambda i, j, ks: np.array([i + 4 * (j + 4 * k) for k in ks])
    spec = [(col(0, 0, [0, 1]), "rate", inj_rate, [0.0, 0.0, 1.0], 1, inj_lims, True),
            (col(3, 3, [0, 1, 2, 3]), prod_type, prod_val, [0.0, 1.0, 0.0], -1, prod_lims, True),
            (col(0, 3, [2, 3]), "bhp", 190 * BARSA, [0.0, 1.0, 0.0], -1, None, False)]
    wo = wd = None
    for cells, kind, val, compi, sign, lims, status in spec:
        wo = og.add_well(wo, go, ro, cells, kind, val, compi=np.array(compi), sign=sign, gravity=grav,
                         lims=None if lims is None else dict(lims), status=status)
        wd = pa.add_well(wd, gd, rd, cells, kind, val, compi=np.array(compi), sign=sign, gravity=grav,
                         lims=None if lims is None else dict(lims), status=status)
    N, T, _, _ = og.internal_connections(go, og.compute_trans(go, ro))
    deck = synthetic_deck()
    mo = GenericBOModel(go, ro, BOFluid(deck), wo, N, T, gravity=og.NORM_GRAVITY)
    md = GenericBlackOilModelAD(gd, rd, init_deck_blackoil_fluid(deck), wd, disgas=True, gravity=og.NORM_GRAVITY,
                                linear_solver=CPRSolver(rtol=1e-12))
    assert isinstance(md, GenericLiveOilModelAD)
    # hydrostatic column, a gas cap in the top layer over undersaturated oil
    so = mo.initial_state(200 * BARSA, 0.25, 0.0)
    z = go.cell_centroids[:, 2]
    so.pressure = 200 * BARSA + 800.0 * og.NORM_GRAVITY * (z - z.min())
    cap = np.arange(80) < 16
    sg = np.where(cap, 0.15, 0.0)
    so.s = np.column_stack([np.full(80, 0.25), 0.75 - sg, sg])
    rsat = value(mo.fluid.rs_sat(so.pressure))
    so.rs = np.where(cap, rsat, 0.85 * rsat)
    so.status = np.where(cap, 3, 1).astype(np.int64)
    so.bhp[0] = 215 * BARSA
    so.qWs[0], so.qOs[0], so.qGs[0] = 0.0, 0.0, inj_rate
    sd = md.initial_state(200 * BARSA, 0.25, 0.0)
    for name in ("pressure", "s", "rs", "status", "qWs", "qOs", "qGs", "bhp"):
        setattr(sd, name, np.array(getattr(so, name), copy=True))
    return mo, so, md, sd


def _compare(mo, so, md, sd, dts):
    from pymrst_amd.nonlinear import NonLinearSolver
    ref_states, ref_its, ref_cuts, _ = simulate_stabilised(mo, so, dts, StabilisedNewton())
    states, rep = NonLinearSolver().simulate_schedule(md, sd, dts)
    print("gpu", [s.iterations for s in rep.steps], md.switches, "oracle", ref_its.tolist(), mo.switches,
          flush=True)
    assert rep.converged and len(ref_states) == len(dts)
    assert [s.iterations for s in rep.steps] == ref_its.tolist()
    assert [s.cuts for s in rep.steps] == ref_cuts.tolist()
    assert md.switches == mo.switches
    assert [w.type for w in md.wells] == [w.type for w in mo.wells]
    for a, b in zip(states, ref_states):
        a = a.to_host() if hasattr(a, "to_host") else a
        assert np.array_equal(a.status, b.status)
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.bhp, b.bhp, rtol=1e-8)
        np.testing.assert_allclose(a.s, b.s, rtol=0, atol=1e-8)
        np.testing.assert_allclose(a.rs, b.rs, rtol=0, atol=1e-8 * np.abs(b.rs).max())
        for q, qb in ((a.qWs, b.qWs), (a.qOs, b.qOs), (a.qGs, b.qGs)):
            np.testing.assert_allclose(q, qb, rtol=1e-7, atol=1e-12)
        for fa, fb in zip(a.well_flux, b.well_flux):
            np.testing.assert_allclose(fa, fb, rtol=1e-7, atol=1e-14)
    return states, ref_states


def test_live_oil_producer_switches_to_its_oil_rate_limit():
    mo, so, md, sd = _models(prod_lims={"orat": -2e-4})
    _compare(mo, so, md, sd, [2 * DAY] * 3)
    assert mo.switches == [(1, "bhp", "orat")]


def test_live_oil_producer_on_a_gas_rate_target():
    mo, so, md, sd = _models(prod_type="grat", prod_val=-0.02, prod_lims={"bhp": 185 * BARSA})
    states, _ = _compare(mo, so, md, sd, [2 * DAY] * 3)
    last = states[-1].to_host() if hasattr(states[-1], "to_host") else states[-1]
    assert last.qGs[1] == pytest.approx(-0.02, rel=1e-9)


def test_live_oil_gas_injector_switches_to_its_bhp_limit():
    mo, so, md, sd = _models(inj_rate=0.3, inj_lims={"bhp": 205 * BARSA})
    states, ref = _compare(mo, so, md, sd, [2 * DAY] * 3)
    assert mo.switches == [(0, "rate", "bhp")]
    assert int((ref[-1].status == 3).sum()) > int((so.status == 3).sum())    # gas came out of solution
