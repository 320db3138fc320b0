"""GPU parity of GenericBlackOilModelAD's oil-water form with the
FacilityModel wells (pymrst_amd.generic) against the oracle restatement
(oracle.generic.GenericOWModel), both driven by the same Newton protocol
(NonLinearSolver / StabilisedNewton), through the C ABI (pmx_ow_set_controls,
pmx_ow_convergence_current_pv, pmx_ow_perf_volume_rates).

Cases: the reference's generic-model self-tests (a rate injector that
switches to its bhp limit, a bhp producer that switches to its liquid-rate
limit; generic_black_oil_model.py:1015-1250), and a gravity column with
multi-perforation wells whose wellbore column is carried from step to step
in the perforation fluxes, producers whose top perforations inject, one of
them switching to its liquid-rate limit, and a shut well.  Bar: the same Newton
iterations and cuts per report step, the same control switches in the same
order, pressure within 1e-8 relative, saturations and perforation fluxes
within 1e-8.
"""
import numpy as np
import pytest

import oracle.grid as og
from oracle.generic import GenericOWModel
from oracle.nonlinear import StabilisedNewton, simulate_stabilised
from oracle.ow import OWFluid

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

BARSA, DAY = og.BARSA, og.DAY
MD = 1e-3 * 9.869233e-13
CP = 1e-3
MU, RHO, NEXP = [1.0 * CP, 5.0 * CP], [1000.0, 800.0], [2.0, 2.0]


def _models(dims, phys, perm, poro, wspec, c, c_r, gravity, p0, sw0, p_init=None):
    """The same case as an oracle model and a device model.  wspec rows:
    (cells, control, value, compi, sign, lims, status)."""
    import pymrst_amd as pa
    from pymrst_amd import ow
    from pymrst_amd.generic import GenericBlackOilModelAD
    grav = np.array([0, 0, gravity]) if gravity else None
    go = og.cart_grid(dims, phys)
    ro = og.make_rock(go, perm, poro)
    gd = pa.cart_grid(dims, phys)
    rd = pa.make_rock(gd, perm, poro)
    wo = wd = None
    for cells, kind, val, compi, sign, lims, status in wspec:
        wo = og.add_well(wo, go, ro, cells, kind, val, compi=np.array(compi, float), sign=sign, gravity=grav,
                         lims=None if lims is None else dict(lims), status=status)
        wd = pa.add_well(wd, gd, rd, cells, kind, val, compi=np.array(compi, float), sign=sign, gravity=grav,
                         lims=None if lims is None else dict(lims), status=status)
    N, T, _, _ = og.internal_connections(go, og.compute_trans(go, ro))
    fo = OWFluid(mu=np.array(MU), rhoS=np.array(RHO), n=np.array(NEXP), c=np.asarray(c, float), p_ref=200 * BARSA,
                 c_r=c_r)
    mo = GenericOWModel(go, ro, fo, wo, N, T, gravity=gravity)
    fd = ow.init_simple_adi_fluid(MU, RHO, NEXP, c, 200 * BARSA, c_r=c_r)
    md = GenericBlackOilModelAD(gd, rd, fd, wd, gravity=gravity, linear_solver=ow.CPRSolver(rtol=1e-12))
    so = mo.initial_state(p0, sw0)
    sd = md.initial_state(p0, sw0)
    if p_init is not None:
        so.pressure = p_init(go.cell_centroids[:, 2])
        sd.pressure = so.pressure.copy()
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
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.sW, b.sW, rtol=0, atol=1e-8)
        np.testing.assert_allclose(a.bhp, b.bhp, rtol=1e-8)
        for q, qo in ((a.qWs, b.qWs), (a.qOs, b.qOs)):
            np.testing.assert_allclose(q, qo, rtol=1e-7, atol=1e-12)
        assert a.well_flux is not None and b.well_flux is not None
        for fa, fb in zip(a.well_flux, b.well_flux):
            np.testing.assert_allclose(fa, fb, rtol=1e-7, atol=1e-12)
    return states, ref_states


def _eight_by_eight(wspec):
    perm = np.full((64, 1), 100 * MD)
    return _models([8, 8, 1], [200.0, 200.0, 10.0], perm, np.full(64, 0.25), wspec,
                   [1e-6 / BARSA, 1e-4 / BARSA], None, 0.0, 200 * BARSA, 0.2)


def test_injector_switches_to_its_bhp_limit():
    p_lim = 250.0 * BARSA
    mo, so, md, sd = _eight_by_eight([([0], "rate", 5e-3, [1, 0], 1, {"bhp": p_lim}, True),
                                      ([63], "bhp", 150 * BARSA, [0, 1], -1, None, True)])
    states, _ = _compare(mo, so, md, sd, np.full(2, 30 * DAY))
    assert md.switches == [(0, "rate", "bhp")]
    np.testing.assert_allclose(states[-1].bhp[0], p_lim, rtol=1e-12)


def test_producer_switches_to_its_liquid_rate_limit():
    l_lim = -5e-4
    mo, so, md, sd = _eight_by_eight([([0], "rate", 2e-3, [1, 0], 1, None, True),
                                      ([63], "bhp", 100 * BARSA, [0, 1], -1, {"lrat": l_lim}, True)])
    states, _ = _compare(mo, so, md, sd, np.full(2, 30 * DAY))
    assert md.switches == [(1, "bhp", "lrat")]
    np.testing.assert_allclose(states[-1].qWs[1] + states[-1].qOs[1], l_lim, rtol=1e-8)


def test_gravity_column_with_cross_flow_switch_and_shut_well():
    """A 4x4x5 column with gravity, rock compressibility and kz = kx / 100,
    started from a vertical pressure gradient of 0.25 bar/m (steeper than
    hydrostatic): a rate injector in three cells with a bhp limit, two bhp
    producers whose top perforations inject while the deeper ones produce
    (one reaches its liquid-rate limit), a shut well.  The column drops are
    taken from each converged step's perforation fluxes."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 4, 4, 5
    perm = np.exp(rng.normal(np.log(200 * MD), 0.5, (nx * ny * nz, 1))) * np.array([[1.0, 1.0, 0.01]])
    col = lambda i, j, ks: np.array([i + nx * (j + ny * k) for k in ks])
    wspec = [(col(0, 0, [2, 3, 4]), "rate", 2e-5, [1, 0], 1, {"bhp": 200.9 * BARSA}, True),
             (col(3, 3, [0, 1, 2, 3]), "bhp", 198 * BARSA, [0, 1], -1, {"lrat": -1.5e-5}, True),
             (col(0, 3, [0, 1, 2, 3, 4]), "bhp", 198 * BARSA, [0, 1], -1, None, True),
             (col(3, 0, [1, 2]), "bhp", 190 * BARSA, [0, 1], -1, None, False)]
    mo, so, md, sd = _models([nx, ny, nz], [80.0, 80.0, 25.0], perm, np.full(nx * ny * nz, 0.2), wspec,
                             [1e-6 / BARSA, 1e-4 / BARSA], 1e-10 / BARSA, og.NORM_GRAVITY, 200 * BARSA, 0.3,
                             p_init=lambda z: 200 * BARSA + 0.25e5 * (z - 12.5))
    states, ref = _compare(mo, so, md, sd, np.array([0.5, 2.0, 5.0, 10.0]) * DAY)
    assert md.switches == [(1, "bhp", "lrat")]
    for s in states:
        assert abs(s.qWs[3]) <= 1e-14 and abs(s.qOs[3]) <= 1e-14
        assert not s.well_flux[3].any()
    # some producer perforation injected: the columns were not single streams
    assert any((s.well_flux[k].sum(axis=1) > 0).any() for s in ref for k in (1, 2))
    np.testing.assert_allclose(np.concatenate(md.last_cdp), np.concatenate(mo.compute_cdp_per_well(ref[-2])),
                               rtol=1e-10)


def test_generic_model_reproduces_the_oil_water_model_on_the_device():
    """The reference's own proof, on the device: without limits the generic
    model and the hand-rolled TwoPhaseOilWaterModelAD step to the same
    states (generic_black_oil_model.py:1060-1120)."""
    import pymrst_amd as pa
    from pymrst_amd import ow
    from pymrst_amd.generic import GenericBlackOilModelAD
    from pymrst_amd.nonlinear import NonLinearSolver
    g = pa.cart_grid([8, 8, 1], [200.0, 200.0, 10.0])
    rock = pa.make_rock(g, np.full((64, 1), 100 * MD), np.full(64, 0.25))
    fl = ow.init_simple_adi_fluid(MU, RHO, NEXP, [1e-6 / BARSA, 1e-4 / BARSA], 200 * BARSA)

    def wells():
        w = pa.add_well(None, g, rock, [0], "rate", 1e-3, compi=np.array([1.0, 0.0]), sign=1)
        return pa.add_well(w, g, rock, [63], "bhp", 150 * BARSA, compi=np.array([0.0, 1.0]), sign=-1)

    tight = dict(tol_cnv=1e-8, tol_mb=1e-12, tol_well=1e-10)
    mg = GenericBlackOilModelAD(g, rock, fl, wells(), linear_solver=ow.CPRSolver(rtol=1e-12), **tight)
    mh = ow.TwoPhaseOilWaterModelAD(g, rock, fl, wells(), linear_solver=ow.CPRSolver(rtol=1e-12))
    mh.tol_cnv, mh.tol_mb, mh.tol_well = tight["tol_cnv"], tight["tol_mb"], tight["tol_well"]
    dts = np.full(3, 30 * DAY)
    sg, rg = NonLinearSolver().simulate_schedule(mg, mg.initial_state(200 * BARSA, 0.2), dts)
    sh, rh = NonLinearSolver().simulate_schedule(mh, mh.initial_state(200 * BARSA, 0.2), dts)
    assert rg.converged and rh.converged
    for a, b in zip(sg, sh):
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-10)
        np.testing.assert_allclose(a.sW, b.sW, rtol=0, atol=1e-10)
