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
