"""Launch-shape autotuning: every candidate shape of an AMG level (sweeps
from the ELL copy; CSR with 4, 8 or 16 lanes per row; the same with the
folded launches) gives the preconditioner application of the default shape
to rounding, every shape of the Gram-Schmidt projections gives the default's
solve, and an autotuned run gives the untuned run's Newton iterations and
states, with every level tuned and the choices kept in the process-wide
table for the next system of the same size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _shapes(md):
    from pymrst_amd import _lib
    out = np.zeros(256, np.int32)
    n = ctypes.c_int32()
    _lib.check(md.lib.pmx_ow_launch_shapes(md.h, _lib.ptr(out), out.size, ctypes.byref(n)), "pmx_ow_launch_shapes")
    out = out[:n.value]
    return int(out[0]), out[2:2 + 5 * int(out[1])].reshape(-1, 5)


def _entries(lib):
    from pymrst_amd import _lib
    n = ctypes.c_int32()
    _lib.check(lib.pmx_autotune_entries(ctypes.byref(n)), "pmx_autotune_entries")
    return n.value


@pytest.mark.parametrize("prec", [64, 32])
def test_every_level_shape_gives_the_same_application(prec):
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    from pymrst_amd.workloads import build_model, field_waterflood
    # 72,000 cells: the finest level keeps an ELL copy as a tuning candidate
    case = field_waterflood(dims=(60, 60, 20), steps=1, step_days=10.0, inj_rate=2e-3)
    solver = CPRSolver(precision=prec, use_graphs=False, autotune=True, maxiter=1, local_sweeps=2)
    md, s0 = build_model(case, solver)
    dev = torch.device("cuda")
    d0 = DeviceOWState.from_host(s0, dev)
    d = d0.copy()
    d.p += 3e5 * torch.rand(d.p.shape, dtype=torch.float64, device=dev,
                            generator=torch.Generator(device=dev).manual_seed(1))
    dt = float(case.dts[0])
    md.assemble_device(d0, d, dt, md._cdp_dev)
    _, pmean = md.converged(dt, d)
    n = 2 * md.nc + 3 * md.nw
    r = torch.as_tensor(np.random.default_rng(3).normal(size=n), device=dev)
    z = {}
    for shape in [2, 0, 1, 3, 4, 5, 6]:
        solver.level_shape = shape
        info = _lib.SolveInfo()
        print(f"level shape {shape}, FP{prec} preconditioner", flush=True)
        rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(solver.options()), pmean, _lib.ptr(md._dx), ctypes.byref(info))
        assert rc in (0, 1)
        spmv, lv = _shapes(md)
        assert lv.shape[0] >= 2
        if shape == 0:
            assert lv[0, 1] == 1, "the finest level sweeps from its ELL copy"
        else:
            assert np.all(lv[:, 1] == 0) and np.all(lv[:, 2] == (4 << ((shape - 1) % 3)))
            assert np.all(lv[:, 3] == (1 if shape >= 4 else 0))
        zk = torch.empty_like(r)
        _lib.check(md.lib.pmx_ow_precondition(md.h, ctypes.byref(solver.options()), _lib.ptr(r), _lib.ptr(zk)),
                   "pmx_ow_precondition")
        z[shape] = zk.cpu().numpy()
    ref = z[2]
    tol = 1e-11 if prec == 64 else 2e-5
    for shape, zk in z.items():
        err = np.abs(zk - ref).max() / np.abs(ref).max()
        assert err <= tol, f"shape {shape}: application differs by {err:.2e}"


def test_autotuned_run_matches_the_untuned_run():
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.workloads import build_model, field_waterflood
    case = field_waterflood(dims=(40, 40, 20), steps=2, step_days=15.0, inj_rate=2e-3)
    runs = {}
    for tune in (False, True):
        solver = CPRSolver(precision=32, autotune=tune, local_sweeps=5)
        md, s0 = build_model(case, solver)
        if tune:
            _lib.check(md.lib.pmx_autotune_reset(), "pmx_autotune_reset")
        print(f"autotune={tune}", flush=True)
        st, its = simulate_schedule_ad(s0, md, case.dts, keep_device=True)
        runs[tune] = (st[-1].p.cpu().numpy(), st[-1].sw.cpu().numpy(), list(its), md)
    p0, s0_, its0, _ = runs[False]
    p1, s1, its1, md = runs[True]
    assert its1 == its0
    assert np.abs(p1 - p0).max() <= 1e-7 * np.abs(p0).max()
    assert np.abs(s1 - s0_).max() <= 1e-7
    spmv, lv = _shapes(md)
    assert spmv in (4, 8, 16)
    assert np.all(lv[:, 4] == 1), "every level above the coarse cycle is tuned"
    assert np.all(np.isin(lv[:, 2], (4, 8, 16)))
    assert md.launch_shapes()["projection_elems"] in (2, 4, 8, 104, 108)
    entries = _entries(md.lib)
    assert entries >= 1 + lv.shape[0]
    # a second model of the same size reuses the table: no new entries
    solver = CPRSolver(precision=32, autotune=True, local_sweeps=5)
    m2, s2 = build_model(case, solver)
    simulate_schedule_ad(s2, m2, case.dts[:1], keep_device=True)
    assert _entries(m2.lib) == entries
    spmv2, lv2 = _shapes(m2)
    assert spmv2 == spmv and np.array_equal(lv2[:, 1:4], lv[:, 1:4])


def test_every_projection_shape_gives_the_same_solve():
    """The Gram-Schmidt projections in each shape (2, 4, 8 elements per
    thread with shuffle sums; 4 or 8 with the sums through LDS) give the
    default shape's FGMRES solve: the same iteration count within one and
    the same increment to the solve's tolerance."""
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    from pymrst_amd.workloads import build_model, field_waterflood
    case = field_waterflood(dims=(40, 40, 20), steps=1, step_days=10.0, inj_rate=2e-3)
    solver = CPRSolver(precision=32, rtol=1e-10, local_sweeps=5)
    md, s0 = build_model(case, solver)
    dev = torch.device("cuda")
    d0 = DeviceOWState.from_host(s0, dev)
    d = d0.copy()
    d.p += 3e5 * torch.rand(d.p.shape, dtype=torch.float64, device=dev,
                            generator=torch.Generator(device=dev).manual_seed(2))
    dt = float(case.dts[0])
    md.assemble_device(d0, d, dt, md._cdp_dev)
    _, pmean = md.converged(dt, d)
    out = {}
    for shape in [8, 2, 4, 104, 108]:
        solver.dot_shape = shape
        info = _lib.SolveInfo()
        print(f"projection shape {shape}", flush=True)
        rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(solver.options()), pmean, _lib.ptr(md._dx), ctypes.byref(info))
        assert rc == 0
        assert md.launch_shapes()["projection_elems"] == shape
        out[shape] = (int(info.iterations), md._dx.cpu().numpy().copy())
    its0, dx0 = out[8]
    for shape, (its, dx) in out.items():
        assert abs(its - its0) <= 1, f"shape {shape}: {its} iterations against {its0}"
        assert np.abs(dx - dx0).max() <= 1e-7 * np.abs(dx0).max(), f"shape {shape}"
