"""Discrete adjoint on the GPU: the transposed CPR solve against a direct
solve of J^T, and log-permeability and control gradients against central
finite differences of the full nonlinear re-simulation (the reference's own
check, pymrst/modules/optimization/adjoint.py:437-469)."""
import ctypes

import numpy as np
import pytest

import oracle.grid as og

from cases import build_device, build_oracle, device_to_reference_perm, small_gravity_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_transposed_solve_matches_direct():
    """(dR/dx)^T lam = b by CPR-FGMRES on the transposed device system equals
    scipy's direct solve of J^T (the oracle's Jacobian) to 1e-7."""
    import scipy.sparse.linalg as spla
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    inp = small_gravity_inputs()
    mo, so = build_oracle(inp)
    rng = np.random.default_rng(2)
    st = so.copy()
    st.pressure = st.pressure + rng.normal(0, 2e5, st.pressure.size)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, st.sW.size), 0.0, 1.0)
    cdp = mo.compute_cdp_per_well(so)
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, 10 * og.DAY, cdp)
    perm = device_to_reference_perm(mo.nc, len(mo.wells))
    lam_true = rng.standard_normal(J_ref.shape[0])
    b_ref = J_ref.T @ lam_true
    lam_ref = spla.spsolve(J_ref.T.tocsc(), b_ref)
    for prec in (64, 32):
        md, sd = build_device(inp, CPRSolver(rtol=1e-10, precision=prec))
        d0, d = DeviceOWState.from_host(so, md.device), DeviceOWState.from_host(st, md.device)
        md.assemble_device(d0, d, 10 * og.DAY, torch.as_tensor(np.concatenate(cdp), device="cuda"))
        ok, pmean = md.converged(10 * og.DAY, d)
        b = torch.as_tensor(b_ref[perm], device="cuda")
        lam = torch.empty_like(b)
        info = _lib.SolveInfo()
        rc = md.lib.pmx_ow_adjoint_solve(md.h, ctypes.byref(md.linear_solver.options()), pmean, _lib.ptr(b),
                                         _lib.ptr(lam), ctypes.byref(info))
        assert rc == 0, (f"FP{prec}: adjoint solve not converged ({info.iterations} its, "
                         f"rel residual {info.rel_residual:.1e})")
        lam_h = lam.cpu().numpy()
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        res = np.linalg.norm(J_ref.T @ lam_h[inv] - b_ref) / np.linalg.norm(b_ref)
        assert res <= 1e-10, f"FP{prec}: ||J^T lam - b|| / ||b|| = {res:.2e}"
        err = float(np.abs(lam_h - lam_ref[perm]).max() / np.abs(lam_ref).max())
        assert err <= 1e-6, f"FP{prec}: transposed CPR differs from the direct solve: {err:.2e}"


def _small_problem(**kw):
    from pymrst_amd import grid as G
    from pymrst_amd.adjoint import AdjointProblem
    from pymrst_amd.ow import CPRSolver, init_simple_adi_fluid
    dims = (8, 7, 3)
    nc = int(np.prod(dims))
    g = G.cart_grid(dims, (8 * 10.0, 7 * 10.0, 3 * 2.0))
    rng = np.random.default_rng(11)
    logk = np.log(100 * og.MILLI * og.DARCY) + 0.8 * rng.standard_normal(nc)
    col = lambda i, j: i + 8 * j + 56 * np.arange(3)
    specs = [dict(cells=col(3, 3), type="rate", val=3e-4, compi=[1.0, 0.0]),
             dict(cells=col(0, 0), type="bhp", val=150 * og.BARSA, compi=[0.0, 1.0]),
             dict(cells=col(7, 6), type="bhp", val=160 * og.BARSA, compi=[0.0, 1.0])]
    fluid = init_simple_adi_fluid((0.3 * og.CENTI * og.POISE, 3.0 * og.CENTI * og.POISE), (1000.0, 850.0),
                                  (2.0, 2.0), (1e-9, 1e-9), 200 * og.BARSA, c_r=1e-10)
    targets = -1e-4 * (1.0 + 0.3 * rng.standard_normal((4, 2)))
    prob = AdjointProblem(g, np.full(nc, 0.2), fluid, specs, np.full(4, 20 * og.DAY), 200 * og.BARSA, 0.2,
                          solver=CPRSolver(rtol=1e-13, precision=64), targets=targets, **kw)
    return prob, logk


def test_logperm_gradient_matches_finite_differences():
    """dJ/dln k from the adjoint equals central differences (h = 1e-6) of
    the re-simulated objective on perforated and interior cells: to 1e-5
    relative to the largest gradient entry, and to 1e-4 cell by cell (the
    finite differences are stable to 1e-8 over h = 1e-4 .. 1e-7; the
    remaining gap is the precision of one transposed solve of a Jacobian
    whose condition number is about 1e11)."""
    prob, logk = _small_problem()
    J, g, gu, run = prob.value_and_gradient(logk)
    g = g.cpu().numpy()
    cells = [3 + 8 * 3, 0, 7 + 8 * 6 + 56, 13, 40, 100, 150, 167]
    h = 1e-6
    fd = np.zeros(len(cells))
    for i, c in enumerate(cells):
        lp, lm = logk.copy(), logk.copy()
        lp[c] += h
        lm[c] -= h
        fd[i] = (prob.value(prob.forward(lp)) - prob.value(prob.forward(lm))) / (2 * h)
    adj = g[cells]
    norm_err = np.abs(adj - fd).max() / np.abs(fd).max()
    assert norm_err <= 1e-5, f"max |adjoint - FD| / max |FD| = {norm_err:.1e}"
    cell_err = np.abs(adj - fd) / np.abs(fd)
    assert cell_err.max() <= 1e-4, f"per-cell relative errors {cell_err}"


def test_control_gradient_matches_finite_differences():
    """dJ/du (the reference's gradient, one control per well per step) for
    the injector rate and a producer bhp, against central differences."""
    prob, logk = _small_problem()
    J, g, gu, run = prob.value_and_gradient(logk)
    # perturb one control over the whole schedule: the gradient is the sum over steps
    for w, h in ((0, 1e-9), (1, 10.0)):
        vals = []
        for sgn in (1, -1):
            prob.well_specs[w]["val"] += sgn * h
            vals.append(prob.value(prob.forward(logk)))
            prob.well_specs[w]["val"] -= sgn * h
        fd = (vals[0] - vals[1]) / (2 * h)
        adj = gu[:, w].sum()
        assert abs(fd - adj) <= 1e-5 * max(abs(fd), 1e-30), f"well {w}: adjoint {adj:.6e} vs FD {fd:.6e}"


def test_checkpointed_sweep_and_autograd_match():
    """Recomputing states between checkpoints (every 3rd ministep) gives the
    same gradient; the torch.autograd.Function returns it as well."""
    from pymrst_amd.adjoint import log_perm_objective
    prob, logk = _small_problem()
    J1, g1, _, _ = prob.value_and_gradient(logk)
    prob3, _ = _small_problem(checkpoint_every=3)
    J3, g3, _, _ = prob3.value_and_gradient(logk)
    assert J1 == J3
    assert torch.allclose(g1, g3, rtol=1e-12, atol=0.0)
    x = torch.as_tensor(logk, dtype=torch.float64, device="cuda").requires_grad_(True)
    J = log_perm_objective(x, prob)
    (gx,) = torch.autograd.grad(J, x)
    assert abs(float(J.detach()) - J1) <= 1e-14 * abs(J1)
    assert torch.allclose(gx, g1, rtol=1e-12, atol=0.0)
