"""GPU parity of the oil-water Newton path against the oracle, through the
C ABI (pmx_ow_*)."""
import numpy as np
import pytest

import oracle.grid as og
from oracle.ow import simulate_schedule

from cases import (build_device, build_oracle, device_to_reference_perm, golden, p5_inputs, paired_inputs,
                   small_gravity_inputs)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev_state(st, device="cuda"):
    from pymrst_amd.ow import DeviceOWState
    return DeviceOWState.from_host(st, torch.device(device))


def _perturbed(s0, seed, wells=True):
    rng = np.random.default_rng(seed)
    st = s0.copy()
    st.pressure = st.pressure + rng.normal(0, 2e5, st.pressure.size)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, st.sW.size), 0.0, 1.0)
    if wells:
        st.qWs = st.qWs + 1e-4
        st.qOs = st.qOs - 2e-4
        st.bhp = st.bhp + 1e5
    return st


@pytest.mark.parametrize("case", ["paired", "gravity"])
def test_assembly_matches_oracle(case):
    """Residual, block values and J x against the oracle's AD Jacobian."""
    inp = paired_inputs() if case == "paired" else small_gravity_inputs()
    mo, so = build_oracle(inp)
    md, sd = build_device(inp)
    dt = 10 * og.DAY
    st = _perturbed(so, 7)
    cdp = mo.compute_cdp_per_well(so)
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, dt, cdp)
    cdp_dev = torch.as_tensor(np.concatenate(cdp), device="cuda")
    md.assemble_device(_dev_state(so), _dev_state(st), dt, cdp_dev)
    nc, nw = mo.nc, len(mo.wells)
    perm = device_to_reference_perm(nc, nw)
    r_dev = md.residual_device().cpu().numpy()
    # residual: scale by the largest accumulation/flux magnitude per equation
    scale = np.abs(r_ref).max()
    assert np.abs(r_dev - r_ref[perm]).max() <= 1e-12 * scale
    # pattern: the TPFA block pattern, bit-exact
    import ctypes
    from pymrst_amd import _lib
    nnzb = ctypes.c_int32()
    _lib.check(md.lib.pmx_ow_nnzb(md.h, ctypes.byref(nnzb)))
    rowptr = np.zeros(nc + 1, np.int32)
    col = np.zeros(nnzb.value, np.int32)
    _lib.check(md.lib.pmx_ow_export_pattern(md.h, _lib.ptr(rowptr), _lib.ptr(col)))
    Acc = J_ref[:nc, :nc] + J_ref[:nc, nc:2 * nc] + J_ref[nc:2 * nc, :nc] + J_ref[nc:2 * nc, nc:2 * nc]
    N = mo.ops.N
    expect = set(zip(N[:, 0], N[:, 1])) | set(zip(N[:, 1], N[:, 0])) | {(i, i) for i in range(nc)}
    got = {(i, int(c)) for i in range(nc) for c in col[rowptr[i]:rowptr[i + 1]]}
    assert got == expect
    assert all(np.all(np.diff(col[rowptr[i]:rowptr[i + 1]]) > 0) for i in range(nc))
    # every nonzero of the oracle's cell-cell blocks lies in the TPFA pattern
    # or couples two perforated cells of one well (mixture terms)
    perf = set(int(c) for w in mo.wells for c in w.cells)
    Acc = Acc.tocoo()
    for i, j in zip(Acc.row, Acc.col):
        assert (i, j) in expect or (i in perf and j in perf)
    # block values vs the oracle at block positions (well mixture couplings
    # between distinct perforated cells are kept outside the blocks)
    vals = torch.empty(nnzb.value * 4, dtype=torch.float64, device="cuda")
    _lib.check(md.lib.pmx_ow_export_blocks(md.h, _lib.ptr(vals)))
    vals = vals.cpu().numpy().reshape(-1, 2, 2)
    Jd = J_ref.tocsr()
    rows = np.repeat(np.arange(nc), np.diff(rowptr))
    ref_blocks = np.stack([np.asarray(Jd[rows, col]).ravel(), np.asarray(Jd[rows, nc + col]).ravel(),
                           np.asarray(Jd[nc + rows, col]).ravel(), np.asarray(Jd[nc + rows, nc + col]).ravel()],
                          axis=1).reshape(-1, 2, 2)
    colscale = np.array([np.abs(J_ref[:, :nc]).max(), np.abs(J_ref[:, nc:2 * nc]).max()])
    err = np.abs(vals - ref_blocks) / colscale[None, None, :]
    mixed = np.array([(r in perf and c in perf and r != c) for r, c in zip(rows, col)])
    has_mix_blocks = mixed.any()
    err_ok = err[~mixed] if has_mix_blocks else err
    assert err_ok.max() <= 1e-12
    # the full operator, mixture terms included: J x for random x
    rng = np.random.default_rng(0)
    for _ in range(3):
        x_ref = rng.normal(size=J_ref.shape[0])
        y_ref = J_ref @ x_ref
        x_dev = torch.as_tensor(x_ref[perm], device="cuda")
        y_dev = torch.empty_like(x_dev)
        _lib.check(md.lib.pmx_ow_matvec(md.h, _lib.ptr(x_dev), _lib.ptr(y_dev)))
        y = y_dev.cpu().numpy()
        ys = np.asarray(abs(J_ref) @ np.abs(x_ref)).ravel()
        assert np.all(np.abs(y - y_ref[perm]) <= 1e-12 * (ys[perm] + 1e-300))


@pytest.mark.parametrize("perturb", [False, True])
def test_linear_solve_matches_direct(perturb):
    """CPR-FGMRES solution equals the direct solve to 1e-7 relative (the
    reference's own contract, linear_solvers.py:639-649, at the previous
    state and at a perturbed one), FP64 and FP32 preconditioners."""
    import scipy.sparse.linalg as spla
    from pymrst_amd.ow import CPRSolver
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    st = _perturbed(so, 11, wells=False) if perturb else so.copy()
    cdp = mo.compute_cdp_per_well(so)
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, 10 * og.DAY, cdp)
    x_ref = spla.spsolve(J_ref.tocsc(), -r_ref)
    perm = device_to_reference_perm(mo.nc, len(mo.wells))
    for prec in (64, 32):
        md, sd = build_device(inp, CPRSolver(rtol=1e-10, precision=prec))
        md.assemble_device(_dev_state(so), _dev_state(st), 10 * og.DAY,
                           torch.as_tensor(np.concatenate(cdp), device="cuda"))
        ok, pmean = md.converged(10 * og.DAY, _dev_state(st))
        import ctypes
        from pymrst_amd import _lib
        info = _lib.SolveInfo()
        rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(md.linear_solver.options()), pmean, _lib.ptr(md._dx),
                                 ctypes.byref(info))
        assert rc == 0, f"FP{prec}: not converged ({info.iterations} its, rel {info.rel_residual:.1e})"
        x = md._dx.cpu().numpy()
        err = float(np.abs(x - x_ref[perm]).max() / np.abs(x_ref).max())
        assert err <= 1e-7, f"FP{prec}: CPR is not the direct solution: {err:.2e} ({info.iterations} its)"


def _run(inp, solver):
    from pymrst_amd.ow import simulate_schedule_ad
    md, s0 = build_device(inp, solver)
    return simulate_schedule_ad(s0, md, inp["dts"])


def test_paired_waterflood_matches_reference_run():
    """Newton counts identical to the reference's run; pressure within 1e-8
    relative, saturation within 1e-8 absolute."""
    from pymrst_amd.ow import CPRSolver
    ref = golden("paired_waterflood.npz")
    states, its = _run(paired_inputs(), CPRSolver(rtol=1e-10))
    assert np.array_equal(its, ref["newton_iterations"]), f"newton its {its.tolist()}"
    p = np.vstack([s.pressure for s in states])
    sw = np.vstack([s.sW for s in states])
    assert np.abs(p - ref["pressure"]).max() <= 1e-8 * np.abs(ref["pressure"]).max()
    assert np.abs(sw - ref["sw"]).max() <= 1e-8
    assert np.abs(np.vstack([s.bhp for s in states]) - ref["bhp"]).max() <= 1e-8 * np.abs(ref["bhp"]).max()


def test_mrst_waterflood_parity():
    from pymrst_amd.ow import CPRSolver
    inp = p5_inputs()
    states, its = _run(inp, CPRSolver(rtol=1e-10))
    mo, so = build_oracle(inp)
    ostates, oits = simulate_schedule(so, mo, inp["dts"])
    assert np.array_equal(its, oits), f"gpu {its.tolist()} oracle {oits.tolist()}"
    ref = golden("p5_ow.npz")
    p = np.vstack([s.pressure for s in states])
    assert np.abs(p - ref["p_hist"]).max() <= 5e-6 * np.abs(ref["p_hist"]).max()
    po = np.vstack([s.pressure for s in ostates])
    assert np.abs(p - po).max() <= 1e-8 * np.abs(po).max()


@pytest.mark.parametrize("prec", [64, 32])
def test_gravity_crossflow_schedule_matches_oracle(prec):
    """Gravity, rock compressibility, a cross-flowing well: same Newton
    counts as the oracle, pressure within 1e-8 (FP64) / 1e-6 (FP32
    preconditioner) relative."""
    from pymrst_amd.ow import CPRSolver
    inp = small_gravity_inputs()
    states, its = _run(inp, CPRSolver(rtol=1e-10, precision=prec))
    mo, so = build_oracle(inp)
    ostates, oits = simulate_schedule(so, mo, inp["dts"])
    assert np.array_equal(its, oits), f"gpu {its.tolist()} oracle {oits.tolist()}"
    p = np.vstack([s.pressure for s in states])
    po = np.vstack([s.pressure for s in ostates])
    assert np.abs(p - po).max() <= 1e-8 * np.abs(po).max()
    sw = np.vstack([s.sW for s in states])
    swo = np.vstack([s.sW for s in ostates])
    assert np.abs(sw - swo).max() <= 1e-8


@pytest.mark.parametrize("prec", [64, 32])
def test_graph_replay_matches_eager(prec):
    """The preconditioner replayed from its captured HIP graph gives the
    eager sequence's result bit for bit, replay after replay, with eager
    work and synchronisations between the replays and a different input
    vector each time."""
    import ctypes
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    md, sd = build_device(inp, CPRSolver(rtol=1e-10, precision=prec))
    st = _perturbed(so, 5, wells=False)
    cdp = torch.as_tensor(np.concatenate(mo.compute_cdp_per_well(so)), device="cuda")
    md.assemble_device(_dev_state(so), _dev_state(st), 10 * og.DAY, cdp)
    ok, pmean = md.converged(10 * og.DAY, _dev_state(st))
    opts = md.linear_solver.options()
    info = _lib.SolveInfo()
    assert md.lib.pmx_ow_solve(md.h, ctypes.byref(opts), pmean, _lib.ptr(md._dx), ctypes.byref(info)) == 0
    n = md._dx.numel()
    gen = torch.Generator(device="cuda").manual_seed(3)
    rs = [torch.randn(n, dtype=torch.float64, device="cuda", generator=gen) for _ in range(3)]

    def apply(r, graphs):
        o = md.linear_solver.options()
        o.use_graphs = int(graphs)
        z = torch.full_like(r, float("nan"))
        _lib.check(md.lib.pmx_ow_precondition(md.h, ctypes.byref(o), _lib.ptr(r), _lib.ptr(z)))
        return z

    eager = [apply(r, False).clone() for r in rs]
    for rep in range(3):
        for k, r in enumerate(rs):
            z = apply(r, True)
            torch.cuda.synchronize()
            assert torch.equal(z, eager[k]), f"replay {rep}, vector {k}: differs from the eager result"
            assert torch.equal(apply(r, False), eager[k])


def test_nonlinear_solver_drives_the_oil_water_model():
    """NonLinearSolver over the oil-water device model (well heads from
    state0 per ministep) equals the model's own step when nothing is cut,
    and the oracle's step when it is."""
    from pymrst_amd.nonlinear import NonLinearSolver
    from pymrst_amd.ow import CPRSolver
    from oracle.nonlinear import simulate_schedule as oracle_schedule
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    md, sd = build_device(inp, CPRSolver(rtol=1e-10))
    dts = np.array([10.0, 30.0]) * og.DAY
    ref_states, ref_its, ref_cuts = oracle_schedule(mo, so, dts)
    states, rep = NonLinearSolver().simulate_schedule(md, sd, dts)
    assert rep.converged and [s.iterations for s in rep.steps] == ref_its.tolist()
    for a, b in zip(states, ref_states):
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.sW, b.sW, rtol=0, atol=1e-8)


@pytest.mark.parametrize("options", [dict(use_relaxation=True),
                                     dict(use_relaxation=True, use_line_search=True, always_use_stabilization=True)])
def test_stabilised_newton_matches_the_oracle(options):
    """Relaxation and the back-tracking line search (with the stabilisation
    forced on) take the oracle's stabilised Newton path: the same Newton
    iterations, cuts and line-search trials per report step, states to 1e-8
    (nonlinear_solver.py:239-381, 596-648)."""
    from pymrst_amd.nonlinear import NonLinearSolver
    from pymrst_amd.ow import CPRSolver
    from oracle.nonlinear import StabilisedNewton, simulate_stabilised
    inp = small_gravity_inputs()
    dts = np.full(3, 30.0 * og.DAY)
    mo, so = build_oracle(inp)
    ref_states, ref_its, ref_cuts, ref_ls = simulate_stabilised(mo, so, dts, StabilisedNewton(**options))
    md, sd = build_device(inp, CPRSolver(rtol=1e-12))
    states, rep = NonLinearSolver(**options).simulate_schedule(md, sd, dts)
    assert rep.converged
    assert [s.iterations for s in rep.steps] == ref_its.tolist()
    assert [s.cuts for s in rep.steps] == ref_cuts.tolist()
    assert [s.line_searches for s in rep.steps] == ref_ls.tolist()
    if options.get("use_line_search"):
        assert ref_ls.sum() > 0          # the line search did back off
    for a, b in zip(states, ref_states):
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.sW, b.sW, rtol=0, atol=1e-8)


@pytest.mark.parametrize("prec", [64, 32])
def test_cpr_solve_on_a_level_large_enough_for_the_ell_smoother(prec):
    """A 66x64x64 field case (270,336 pressure rows: the AMG's first level
    is smoothed from its sliced-ELL copy): one Newton solve of J dx = -r
    reaches the requested true residual, checked with the C-ABI product."""
    import ctypes
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver
    from pymrst_amd.workloads import build_model, field_waterflood
    case = field_waterflood(dims=(66, 64, 64), steps=1)
    md, s0 = build_model(case, CPRSolver(rtol=1e-10, precision=prec))
    ds = md.device_state(s0)
    cdp = md.compute_cdp_per_well(ds)
    md.assemble_device(ds, ds.copy(), float(case.dts[0]), cdp)
    ok, pmean = md.converged(float(case.dts[0]), ds)
    r = md.residual_device()
    info = _lib.SolveInfo()
    rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(md.linear_solver.options()), float(pmean), _lib.ptr(md._dx),
                             ctypes.byref(info))
    assert rc == 0 and info.iterations < 200
    y = torch.empty_like(md._dx)
    _lib.check(md.lib.pmx_ow_matvec(md.h, _lib.ptr(md._dx), _lib.ptr(y)))
    res = torch.linalg.norm(y + r) / torch.linalg.norm(r)
    assert float(res) <= 1e-10 * 1.01, float(res)


def test_selector_driven_schedule_matches_oracle():
    """NonLinearSolver with the iteration-count selector on the oil-water
    device model (gravity, a cross-flowing well) against the oracle's
    selector-driven report steps: identical ministeps and Newton counts,
    pressure within 1e-8 relative, saturation within 1e-8."""
    from pymrst_amd.nonlinear import NonLinearSolver
    from pymrst_amd.ow import CPRSolver
    import oracle.timesteps as ot
    import pymrst_amd.timesteps as pt
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    md, sd = build_device(inp, CPRSolver(rtol=1e-12))
    dts = np.array([20.0, 60.0]) * og.DAY
    kw = dict(target_iteration_count=4, first_rampup_step=1 * og.DAY)
    ref_states, ref_its, ref_mini = ot.simulate_schedule_adaptive(mo, so, dts, ot.IterationCount(**kw))
    states, rep = NonLinearSolver(timestep_selector=pt.IterationCountTimeStepSelectorAD(**kw)).simulate_schedule(
        md, sd, dts)
    assert rep.converged and [s.iterations for s in rep.steps] == ref_its.tolist()
    assert [s.ministeps for s in rep.steps] == ref_mini and sum(len(m) for m in ref_mini) > 2 * len(dts)
    for a, b in zip(states, ref_states):
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.sW, b.sW, rtol=0, atol=1e-8)
