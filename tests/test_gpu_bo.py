"""GPU parity of the three-phase black-oil Newton path against the oracle,
through the C ABI (pmx_bo_*): residual and Jacobian (3x3 blocks, well tail,
wellbore-mixture couplings), the update with the variable switch and flash,
and whole schedules with status switching (same Newton counts, states within
1e-8)."""
import numpy as np
import pytest

import oracle.grid as og
from oracle.bo import BOFluid, BOModel, simulate_schedule, synthetic_deck

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _pair(dims=(4, 3, 3), gravity=og.NORM_GRAVITY, wells="injector", seed=0, linear_solver=None):
    """The same case as an oracle model and a device model (shared
    connections and transmissibilities)."""
    import pymrst_amd as pa
    from pymrst_amd.bo import ThreePhaseBlackOilModelAD, init_deck_blackoil_fluid
    nx, ny, nz = dims
    phys = [30.0 * nx, 30.0 * ny, 5.0 * nz]
    g = og.cart_grid(list(dims), phys)
    rng = np.random.default_rng(seed)
    perm = 100 * og.MILLI * og.DARCY * np.exp(0.5 * rng.standard_normal(g.num_cells))
    poro = 0.2 + 0.05 * rng.random(g.num_cells)
    rock = og.make_rock(g, perm, poro)
    N, T = og.internal_connections(g, og.compute_trans(g, rock))[:2]
    nc = g.num_cells
    ws = []
    if wells in ("injector", "injector1"):
        inj = [0, nx * ny] if wells == "injector" else [0]
        ws = og.add_well(None, g, rock, inj, "rate", 5e-4, compi=[0.0, 0.0, 1.0])
        ws = og.add_well(ws, g, rock, [nc - 1 - nx * ny, nc - 1], "bhp", 150 * og.BARSA, compi=[0.0, 1.0, 0.0])
    elif wells == "crossflow":
        col = [nx * ny * k for k in range(nz)]
        ws = og.add_well(None, g, rock, col, "bhp", 200 * og.BARSA, compi=[0.0, 0.0, 1.0])
        ws = og.add_well(ws, g, rock, [nc - 1 - nx * ny, nc - 1], "rate", -3e-4, compi=[0.0, 1.0, 0.0])
    deck = synthetic_deck()
    mo = BOModel(g, rock, BOFluid(deck), ws, N, T, gravity=gravity)
    gd = pa.cart_grid(list(dims), phys)
    rd = pa.make_rock(gd, perm, poro)
    wd = [pa.Well(cells=np.asarray(w.cells), type=w.type, val=w.val, WI=np.asarray(w.WI, float),
                  dZ=np.zeros(len(w.cells)), compi=np.asarray(w.compi, float)) for w in ws]
    md = ThreePhaseBlackOilModelAD(gd, rd, init_deck_blackoil_fluid(deck), wd, gravity=gravity,
                                   linear_solver=linear_solver, T=T, N=N)
    np.testing.assert_allclose(md.pv_ref, mo.pv_ref, rtol=1e-14)
    return mo, md


def _perm(nc, nw):
    """Device position -> reference (variable-blocked) position."""
    d = np.arange(3 * nc + 4 * nw)
    ref = np.empty_like(d)
    ref[:3 * nc] = (d[:3 * nc] % 3) * nc + d[:3 * nc] // 3
    k = d[3 * nc:] - 3 * nc
    ref[3 * nc:] = 3 * nc + (k % 4) * nw + k // 4
    return ref


def _mixed(m, seed, gas=None, well_p=None):
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
    nw = len(m.wells)
    if nw:
        st.qWs[:] = rng.normal(0, 1e-5, nw)
        st.qOs[:] = -2e-4 + rng.normal(0, 1e-5, nw)
        st.qGs[:] = 4e-4 + rng.normal(0, 1e-5, nw)
        st.bhp[:] = [w.val if w.type == "bhp" else 205 * og.BARSA for w in m.wells]
    if well_p is not None:
        # perforated cells straddle the bhp: some connections inject, some produce
        w = m.wells[0]
        st.pressure[w.cells] = st.bhp[0] + well_p * np.where(np.arange(w.cells.size) % 2 == 0, 1.0, -1.0)
    return st


def _dev(md, st):
    from pymrst_amd.bo import DeviceBOState
    return DeviceBOState.from_host(st, md.device)


@pytest.mark.parametrize("case", ["injector", "crossflow", "nowells"])
def test_assembly_matches_oracle(case):
    """Residual and J x against the oracle's forward-mode AD Jacobian, with
    a random mix of saturated and undersaturated cells; "crossflow" puts the
    perforations of one well on both sides of its bhp, so the wellbore
    mixture of the producing connections drives the injecting ones."""
    mo, md = _pair(wells={"nowells": None}.get(case, case), seed=3)
    st0 = _mixed(mo, 11)
    st = _mixed(mo, 12, gas=st0.status == 3, well_p=(8e5 if case == "crossflow" else None))
    dt = 5 * og.DAY
    r_ref, J_ref, _ = mo.residual_and_jacobian(st0, st, dt)
    md.begin_step(_dev(md, st0))
    md.assemble_device(_dev(md, st), dt)
    nc, nw = mo.nc, len(mo.wells)
    perm = _perm(nc, nw)
    r_dev = md.residual_device().cpu().numpy()
    scale = np.abs(r_ref).max()
    assert np.abs(r_dev - r_ref[perm]).max() <= 1e-12 * scale
    J = J_ref.tocsr()
    Jabs = abs(J)
    rng = np.random.default_rng(5)
    for trial in range(3):
        x_dev = rng.standard_normal(3 * nc + 4 * nw)
        x_ref = np.empty_like(x_dev)
        x_ref[perm] = x_dev
        y_ref = J @ x_ref
        bound = Jabs @ np.abs(x_ref)
        y_dev = md.matvec(torch.as_tensor(x_dev, device="cuda")).cpu().numpy()
        err = np.abs(y_dev - y_ref[perm])
        assert np.all(err <= 1e-12 * bound[perm] + 1e-300), (trial, err.max(), np.argmax(err - 1e-12 * bound[perm]))


def test_update_and_flash_match_oracle():
    """update_state + _flash on a state with large increments: clipped
    saturation changes, negative saturations renormalised, rs capped at the
    saturation curve, gas appearing and disappearing."""
    mo, md = _pair(seed=4)
    nc, nw = mo.nc, len(mo.wells)
    perm = _perm(nc, nw)
    rng = np.random.default_rng(21)
    for trial in range(4):
        st = _mixed(mo, 30 + trial)
        upd = np.concatenate([rng.normal(0, 20e5, nc), rng.normal(0, 0.3, nc),
                              np.where(st.status == 1, rng.normal(0, 40.0, nc), rng.normal(0, 0.3, nc)),
                              rng.normal(0, 1e-4, 3 * nw), rng.normal(0, 1e5, nw)])
        ref = mo.update_state(st, upd)
        ds = _dev(md, st)
        dx = torch.as_tensor(upd[perm], device="cuda")
        from pymrst_amd import _lib
        P = _lib.ptr
        _lib.check(md.lib.pmx_bo_update(md.h, P(dx), md.ds_max, P(ds.p), P(ds.s), P(ds.rs), P(ds.status),
                                        P(ds.q), P(ds.bhp)))
        got = ds.to_host()
        np.testing.assert_array_equal(got.status, ref.status)
        np.testing.assert_allclose(got.pressure, ref.pressure, rtol=1e-15)
        np.testing.assert_allclose(got.s, ref.s, rtol=0, atol=1e-15)
        np.testing.assert_allclose(got.rs, ref.rs, rtol=1e-14, atol=1e-14)
        for a, b in ((got.qWs, ref.qWs), (got.qOs, ref.qOs), (got.qGs, ref.qGs), (got.bhp, ref.bhp)):
            np.testing.assert_allclose(a, b, rtol=1e-15)
        assert (ref.status == 3).any() and (ref.status == 1).any()


@pytest.mark.parametrize("prec", [64, 32])
def test_gas_injection_schedule_matches_oracle(prec):
    """Gas injection into undersaturated oil with a producer below the
    bubble point: cells switch between rs and sG as primary variable.  Same
    Newton counts per step, pressure within 1e-8 relative, saturations within
    1e-8 absolute, rs within 1e-8 relative, identical status.  The injector
    has one perforation: a rate-controlled well that crossflows between two
    perforations under gravity makes the Jacobian nearly singular in the
    uniform (p, bhp) shift (condition ~1e13 after equilibration, against ~1e4
    here), where even two direct solvers disagree beyond 1e-9."""
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    mo, md = _pair(dims=(6, 5, 3), seed=7, wells="injector1",
                   linear_solver=CPRSolver(rtol=1e-12, precision=prec))
    st0 = mo.initial_state(200 * og.BARSA, 0.25, 0.0, rs0=45.0)   # bubble point near 155 bar
    st0.bhp[:] = [200 * og.BARSA, 150 * og.BARSA]
    dts = np.array([1.0, 4.0, 10.0, 20.0, 30.0]) * og.DAY
    ref_states, ref_its = simulate_schedule(st0, mo, dts)
    states, its = simulate_schedule_ad(st0, md, dts)
    np.testing.assert_array_equal(its, ref_its)
    switched = False
    for a, b in zip(states, ref_states):
        np.testing.assert_array_equal(a.status, b.status)
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.s, b.s, rtol=0, atol=1e-8)
        np.testing.assert_allclose(a.rs, b.rs, rtol=1e-8, atol=1e-8)
        np.testing.assert_allclose(a.bhp, b.bhp, rtol=1e-8)
        for x, y in ((a.qWs, b.qWs), (a.qOs, b.qOs), (a.qGs, b.qGs)):
            np.testing.assert_allclose(x, y, rtol=1e-7, atol=1e-12)
        switched |= bool((b.status == 3).any() and (b.status == 1).any())
    assert switched, "the schedule never mixes saturated and undersaturated cells"


@pytest.mark.parametrize("maxit, days", [(4, [30.0, 90.0]), (3, [30.0, 30.0])])
def test_nonlinear_solver_cuts_match_oracle(maxit, days):
    """NonLinearSolver.simulate_schedule on the device model against the
    oracle restatement of solve_timestep / _iter_timesteps: with the Newton
    cap lowered, the first report step is taken as 2^k equal sub-steps
    (same k, same summed Newton count, same states); with a cap too low to
    converge at all, the run stops after the failed report step."""
    from pymrst_amd.nonlinear import NonLinearSolver
    from pymrst_amd.ow import CPRSolver
    from oracle.nonlinear import simulate_schedule as oracle_schedule
    mo, md = _pair(dims=(6, 5, 3), seed=7, wells="injector1", linear_solver=CPRSolver(rtol=1e-12))
    mo.max_iterations = maxit
    st0 = mo.initial_state(200 * og.BARSA, 0.25, 0.0, rs0=45.0)
    st0.bhp[:] = [200 * og.BARSA, 150 * og.BARSA]
    dts = np.array(days) * og.DAY
    ref_states, ref_its, ref_cuts = oracle_schedule(mo, st0, dts)
    events = []
    nls = NonLinearSolver(max_iterations=maxit, observer=lambda e, p: events.append(e))
    states, rep = nls.simulate_schedule(md, st0, dts)
    assert [s.iterations for s in rep.steps] == ref_its.tolist()
    assert [s.cuts for s in rep.steps] == ref_cuts.tolist()
    assert len(states) == len(ref_states)
    assert rep.steps[0].cuts > 0 and "timestep_cut" in events
    for a, b, r in zip(states, ref_states, rep.steps):
        if not r.converged:
            assert events[-1] == "step_rejected"
            continue
        np.testing.assert_array_equal(a.status, b.status)
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.s, b.s, rtol=0, atol=1e-8)


@pytest.mark.parametrize("selector", ["iterations", "state"])
def test_selector_driven_schedule_matches_oracle(selector):
    """NonLinearSolver with a time-step selector (the reference's
    _solve_adaptive) on the device black-oil model against the oracle's
    restatement over the oracle model: the same ministeps (bit for bit when
    they follow from Newton counts; to 1e-9 when the state change sets them),
    the same Newton counts, the same states."""
    from pymrst_amd.nonlinear import NonLinearSolver
    from pymrst_amd.ow import CPRSolver
    import oracle.timesteps as ot
    import pymrst_amd.timesteps as pt
    mo, md = _pair(dims=(6, 5, 3), seed=7, wells="injector1", linear_solver=CPRSolver(rtol=1e-12))
    st0 = mo.initial_state(200 * og.BARSA, 0.25, 0.0, rs0=45.0)
    st0.bhp[:] = [200 * og.BARSA, 150 * og.BARSA]
    dts = np.array([30.0, 90.0]) * og.DAY
    if selector == "iterations":
        kw = dict(target_iteration_count=3, max_timestep=40 * og.DAY, first_rampup_step=2 * og.DAY)
        o, p = ot.IterationCount(**kw), pt.IterationCountTimeStepSelectorAD(**kw)
    else:
        kw = dict(target_props=["pressure", "s"], target_change_abs=[20 * og.BARSA, 0.02],
                  first_rampup_step=2 * og.DAY)
        o, p = ot.StateChange(**kw), pt.StateChangeTimeStepSelectorAD(**kw)
    ref_states, ref_its, ref_mini = ot.simulate_schedule_adaptive(mo, st0, dts, o)
    states, rep = NonLinearSolver(timestep_selector=p).simulate_schedule(md, st0, dts)
    assert rep.converged and [s.iterations for s in rep.steps] == ref_its.tolist()
    assert sum(len(m) for m in ref_mini) > 2 * len(dts), "the selector should subdivide the report steps"
    for r, m in zip(rep.steps, ref_mini):
        if selector == "iterations":
            assert r.ministeps == m
        else:
            np.testing.assert_allclose(r.ministeps, m, rtol=1e-9)
    for a, b in zip(states, ref_states):
        np.testing.assert_array_equal(a.status, b.status)
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.s, b.s, rtol=0, atol=1e-8)
