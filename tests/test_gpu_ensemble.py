"""Batched ensemble: the union solve gives every member the state its own
run gives (tight tolerances so both are converged), and its member-wise
convergence test is not fooled by the other members."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_union_matches_members_run_alone():
    from pymrst_amd.ensemble import build_ensemble
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import build_model, field_waterflood
    cases = [field_waterflood(dims=d, seed=s, steps=2, step_days=10.0, inj_rate=r, sigma=sg)
             for d, s, r, sg in (((7, 6, 3), 5, 2e-4, 0.7), ((5, 9, 4), 6, 3e-4, 1.8), ((8, 4, 2), 7, 1e-4, 1.1))]
    tight = dict(tol_cnv=1e-9, tol_mb=1e-12, tol_well=1e-10)
    model, s0, ranges = build_ensemble(cases, CPRSolver(rtol=1e-12))
    for k, v in tight.items():
        setattr(model, k, v)
    states, its = simulate_schedule_ad(s0, model, cases[0].dts, keep_device=True)
    p_union = states[-1].p.cpu().numpy()
    sw_union = states[-1].sw.cpu().numpy()
    for c, (a, b) in zip(cases, ranges):
        m, st0 = build_model(c, CPRSolver(rtol=1e-12))
        for k, v in tight.items():
            setattr(m, k, v)
        st, _ = simulate_schedule_ad(st0, m, c.dts, keep_device=True)
        p = st[-1].p.cpu().numpy()
        assert np.abs(p_union[a:b] - p).max() <= 1e-8 * np.abs(p).max()
        assert np.abs(sw_union[a:b] - st[-1].sw.cpu().numpy()).max() <= 1e-8


def test_member_convergence_is_per_member():
    """After a converged step, disturbing one member's pressure leaves
    exactly that member unconverged."""
    from pymrst_amd.ensemble import build_ensemble
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import field_waterflood
    cases = [field_waterflood(dims=(6, 5, 3), seed=s, steps=1, step_days=10.0, inj_rate=2e-4) for s in (1, 2, 3)]
    model, s0, ranges = build_ensemble(cases, CPRSolver())
    states, _ = simulate_schedule_ad(s0, model, cases[0].dts, keep_device=True)
    d = states[-1].copy()
    start = model.initial_state(cases[0].p0, cases[0].sw0)
    from pymrst_amd.ow import DeviceOWState
    d0 = DeviceOWState.from_host(start, model.device)
    dt = float(cases[0].dts[0])
    model.assemble_device(d0, d, dt, model._cdp_dev)
    assert model.converged(dt, d)[0]
    a, b = ranges[1]
    d.p[a:b] += 1e3 * torch.linspace(0.0, 1.0, b - a, dtype=torch.float64, device=d.p.device)
    model.assemble_device(d0, d, dt, model._cdp_dev)
    ok, _ = model.converged(dt, d)
    assert not ok and model.unconverged_members == 1


def test_member_convergence_reductions():
    """pmx_ow_member_convergence against the same reductions of each
    member's slice of the assembled residual (members of 126 to 3,000
    cells, so one member spans several reduction chunks): maxima exact,
    sums to rounding; and each member's whole-model reductions
    (pmx_ow_convergence of the member built alone) agree."""
    import ctypes
    from pymrst_amd import _lib
    from pymrst_amd.ensemble import build_ensemble
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    from pymrst_amd.workloads import build_model, field_waterflood
    cases = [field_waterflood(dims=d, seed=s, steps=1, step_days=10.0, inj_rate=2e-4)
             for d, s in (((20, 15, 8), 1), ((7, 6, 3), 2), ((30, 20, 5), 3))]
    model, s0, ranges = build_ensemble(cases, CPRSolver())
    d0 = DeviceOWState.from_host(s0, model.device)
    d = d0.copy()
    g = torch.Generator(device=model.device).manual_seed(4)
    d.p += 5e5 * torch.rand(d.p.shape, dtype=torch.float64, device=model.device, generator=g)
    d.sw += 0.1 * torch.rand(d.sw.shape, dtype=torch.float64, device=model.device, generator=g)
    dt = float(cases[0].dts[0])
    model.assemble_device(d0, d, dt, model._cdp_dev)
    out = np.zeros((3, 7))
    wr = np.zeros(3 * model.nw)
    _lib.check(model.lib.pmx_ow_member_convergence(model.h, _lib.ptr(d.p), _lib.ptr(out), _lib.ptr(wr)),
               "pmx_ow_member_convergence")
    r = model.residual_device().cpu().numpy()
    p = d.p.cpu().numpy()
    f = model.fluid
    for m, (a, b) in enumerate(ranges):
        rw, ro, pv = r[2 * a:2 * b:2], r[2 * a + 1:2 * b:2], model.pv[a:b]
        assert out[m, 0] == np.abs(rw / pv).max()
        assert out[m, 1] == np.abs(ro / pv).max()
        np.testing.assert_allclose(out[m, 2:4], [rw.sum(), ro.sum()], rtol=1e-11, atol=1e-12 * np.abs(r).max())
        ib = [np.exp(-c * (p[a:b] - f.p_ref)).sum() if c else float(b - a) for c in (f.c[0], f.c[1])]
        np.testing.assert_allclose(out[m, 4:6], ib, rtol=1e-13)
        np.testing.assert_allclose(out[m, 6], p[a:b].sum(), rtol=1e-13)
    np.testing.assert_array_equal(wr, r[2 * model.nc:])
    # each member alone, at the same states
    for m, (c, (a, b)) in enumerate(zip(cases, ranges)):
        mm, st0 = build_model(c, CPRSolver())
        e0 = DeviceOWState.from_host(st0, mm.device)
        e = e0.copy()
        e.p.copy_(d.p[a:b])
        e.sw.copy_(d.sw[a:b])
        mm.assemble_device(e0, e, dt, mm._cdp_dev)
        out7, _ = mm.convergence_values(e)
        np.testing.assert_allclose(out[m, :2], out7[:2], rtol=1e-12)
        np.testing.assert_allclose(out[m, 4:7], out7[4:7], rtol=1e-12)


@pytest.mark.parametrize("compact", [0.0, 0.9])
def test_per_member_time_stepping_matches_members_run_alone(compact):
    """simulate_ensemble: every member takes the ministeps, cuts and Newton
    iterations it takes alone under simulate_schedule_ad (a Newton cap of 8
    makes every member cut its first 40-day step, each a different number of
    times), and ends in the same state; also when the members still
    iterating are moved into smaller unions as the others finish
    (compact = 0.9: as soon as one member is done)."""
    from pymrst_amd.ensemble import build_ensemble, simulate_ensemble
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import build_model, field_waterflood
    cases = [field_waterflood(dims=d, seed=s, steps=2, step_days=40.0, inj_rate=r, sigma=sg)
             for d, s, r, sg in (((7, 6, 3), 5, 4e-4, 0.7), ((5, 9, 4), 6, 6e-4, 1.8), ((8, 4, 2), 7, 1e-4, 1.1),
                                 ((6, 6, 3), 8, 8e-4, 1.5))]
    maxit = 8
    model, s0, ranges = build_ensemble(cases, CPRSolver(rtol=1e-12))
    model.max_iterations = maxit
    states, union_its, member_its = simulate_ensemble(model, s0, cases[0].dts, compact=compact)
    cut_somewhere = False
    for k, (c, (a, b)) in enumerate(zip(cases, ranges)):
        m, st0 = build_model(c, CPRSolver(rtol=1e-12))
        m.max_iterations = maxit
        events = []
        st, its = simulate_schedule_ad(st0, m, c.dts, observer=lambda e, p: events.append(e))
        cut_somewhere |= "timestep_cut" in events
        assert member_its[:, k].tolist() == its.tolist(), (k, member_its[:, k], its)
        for step in range(len(c.dts)):
            p_alone = st[step].pressure
            np.testing.assert_allclose(states[step].pressure[a:b], p_alone, rtol=1e-8)
            np.testing.assert_allclose(states[step].sW[a:b], st[step].sW, rtol=0, atol=1e-8)
    assert cut_somewhere
    # the union iterates as long as its slowest member, not the sum of all members' cuts
    assert (union_its <= member_its.sum(axis=1)).all()
