"""The line local stage on irregular lines: the caller's lines
(pmx_ow_set_lines) cut every vertical column into one to three pieces, so
lines of lengths 1 to 5 share the waves of each colour.  Each line is
solved by the twisted factorization (its top half eliminated downwards and
its bottom half upwards, meeting at its middle cell; a line of one cell is
its middle only).  The CPR-FGMRES solution of a perturbed Newton system
equals the direct solve to 1e-7 relative (the reference's CPR contract,
linear_solvers.py:639-649), FP64 and FP32 preconditioners."""
import ctypes

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle.grid as og

from cases import build_device, build_oracle, device_to_reference_perm, small_gravity_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _irregular_lines(nx, ny, nz, seed=3):
    rng = np.random.default_rng(seed)
    lines = []
    for j in range(ny):
        for i in range(nx):
            col = np.array([i + nx * (j + ny * k) for k in range(nz)])
            cuts = np.sort(rng.choice(np.arange(1, nz), size=int(rng.integers(0, 3)), replace=False))
            lines += [seg for seg in np.split(col, cuts)]
    return lines


@pytest.mark.parametrize("prec", [64, 32])
def test_irregular_lines_solve_matches_direct(prec):
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    nx, ny, nz = 8, 6, 5
    inp = small_gravity_inputs(nx=nx, ny=ny, nz=nz)
    mo, so = build_oracle(inp)
    st = so.copy()
    rng = np.random.default_rng(11)
    st.pressure = st.pressure + rng.normal(0, 2e5, mo.nc)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, mo.nc), 0.0, 1.0)
    dt = 10 * og.DAY
    cdp = mo.compute_cdp_per_well(so)
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, dt, cdp)
    x_ref = spla.spsolve(J_ref.tocsc(), -r_ref)
    perm = device_to_reference_perm(mo.nc, len(mo.wells))
    lines = _irregular_lines(nx, ny, nz)
    lengths = [len(seg) for seg in lines]
    assert min(lengths) == 1 and max(lengths) == nz and len(set(lengths)) >= 4
    ptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum(lengths)]), dtype=np.int32)
    cells = np.ascontiguousarray(np.concatenate(lines), dtype=np.int32)
    md, _ = build_device(inp, CPRSolver(rtol=1e-10, precision=prec))
    _lib.check(md.lib.pmx_ow_set_lines(md.h, len(lines), _lib.ptr(ptr), _lib.ptr(cells)), "pmx_ow_set_lines")
    dev = torch.device("cuda")
    md.assemble_device(DeviceOWState.from_host(so, dev), DeviceOWState.from_host(st, dev), dt,
                       torch.as_tensor(np.concatenate(cdp), device=dev))
    ok, pmean = md.converged(dt, DeviceOWState.from_host(st, dev))
    info = _lib.SolveInfo()
    rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(md.linear_solver.options()), pmean, _lib.ptr(md._dx),
                             ctypes.byref(info))
    assert rc == 0, f"FP{prec}: not converged ({info.iterations} its, rel {info.rel_residual:.1e})"
    x = md._dx.cpu().numpy()
    err = float(np.abs(x - x_ref[perm]).max() / np.abs(x_ref).max())
    assert err <= 1e-7, f"FP{prec}: CPR is not the direct solution: {err:.2e} ({info.iterations} its)"
