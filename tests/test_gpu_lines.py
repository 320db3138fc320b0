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


def _coloured_line_gauss_seidel(rowptr, col, blocks, lines, r_cells, sweeps):
    """Block line Gauss-Seidel on the cell blocks: lines ordered longest
    first (stably), each coloured greedily with the smallest colour none of
    its already-coloured neighbouring lines has; then, sweep after sweep, the
    colours in order, each line solved exactly (dense) against the newest
    values of the cells off the line, starting from zero."""
    nc = rowptr.size - 1
    line_of = np.empty(nc, dtype=np.int64)
    for l, seg in enumerate(lines):
        line_of[seg] = l
    order = sorted(range(len(lines)), key=lambda l: -len(lines[l]))
    colour = {}
    for l in order:
        used = {colour[m] for c in lines[l] for m in line_of[col[rowptr[c]:rowptr[c + 1]]]
                if m != l and m in colour}
        pick = 0
        while pick in used:
            pick += 1
        colour[l] = pick
    y = np.zeros((nc, 2))
    for _ in range(sweeps):
        for c in range(max(colour.values()) + 1):
            for l in (l for l in order if colour[l] == c):
                seg = list(lines[l])
                at = {cell: i for i, cell in enumerate(seg)}
                A = np.zeros((2 * len(seg), 2 * len(seg)))
                rhs = r_cells[seg].copy()
                for i, cell in enumerate(seg):
                    for k in range(rowptr[cell], rowptr[cell + 1]):
                        j = col[k]
                        if j in at:
                            A[2 * i:2 * i + 2, 2 * at[j]:2 * at[j] + 2] = blocks[k]
                        else:
                            rhs[i] -= blocks[k] @ y[j]
                y[seg] = np.linalg.solve(A, rhs.ravel()).reshape(-1, 2)
    return y


@pytest.mark.parametrize("sweeps", [1, 2])
def test_line_stage_application_matches_coloured_line_gauss_seidel(sweeps):
    """One application of the line local stage alone (no pressure stage,
    FP64, no graph replay) on irregular lines of lengths 1 to 5 equals a
    dense restatement of coloured block line Gauss-Seidel on the exported
    Jacobian blocks to 1e-11, and each well's unknowns equal the solve with
    its own 3 x 3 block.  This pins the twisted factorization and solve,
    the line colouring and the off-line right-hand sides directly, not
    through an outer Krylov solve that would reach the same answer with a
    wrong preconditioner."""
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    nx, ny, nz = 8, 6, 5
    inp = small_gravity_inputs(nx=nx, ny=ny, nz=nz)
    mo, so = build_oracle(inp)
    st = so.copy()
    rng = np.random.default_rng(5)
    st.pressure = st.pressure + rng.normal(0, 2e5, mo.nc)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, mo.nc), 0.0, 1.0)
    dt = 10 * og.DAY
    cdp = mo.compute_cdp_per_well(so)
    lines = _irregular_lines(nx, ny, nz)
    ptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum([len(s) for s in lines])]), dtype=np.int32)
    cells = np.ascontiguousarray(np.concatenate(lines), dtype=np.int32)
    solver = CPRSolver(rtol=1e-10, precision=64, local_sweeps=sweeps, use_pressure_stage=False, use_graphs=False)
    md, _ = build_device(inp, solver)
    _lib.check(md.lib.pmx_ow_set_lines(md.h, len(lines), _lib.ptr(ptr), _lib.ptr(cells)), "pmx_ow_set_lines")
    dev = torch.device("cuda")
    md.assemble_device(DeviceOWState.from_host(so, dev), DeviceOWState.from_host(st, dev), dt,
                       torch.as_tensor(np.concatenate(cdp), device=dev))
    ok, pmean = md.converged(dt, DeviceOWState.from_host(st, dev))
    opts = solver.options()
    info = _lib.SolveInfo()
    rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(opts), pmean, _lib.ptr(md._dx), ctypes.byref(info))
    assert rc in (0, 1), rc     # the solve sets the preconditioner up; its answer is not used here
    nc, nw = mo.nc, len(mo.wells)
    r = torch.as_tensor(np.random.default_rng(9).normal(size=2 * nc + 3 * nw), device=dev)
    z = torch.empty_like(r)
    print(f"line stage application: {len(lines)} lines, {sweeps} sweeps", flush=True)
    _lib.check(md.lib.pmx_ow_precondition(md.h, ctypes.byref(opts), _lib.ptr(r), _lib.ptr(z)),
               "pmx_ow_precondition")
    nnzb = ctypes.c_int32()
    _lib.check(md.lib.pmx_ow_nnzb(md.h, ctypes.byref(nnzb)))
    rowptr = np.zeros(nc + 1, np.int32)
    col = np.zeros(nnzb.value, np.int32)
    _lib.check(md.lib.pmx_ow_export_pattern(md.h, _lib.ptr(rowptr), _lib.ptr(col)))
    vals = torch.empty(nnzb.value * 4, dtype=torch.float64, device=dev)
    _lib.check(md.lib.pmx_ow_export_blocks(md.h, _lib.ptr(vals)))
    blocks = vals.cpu().numpy().reshape(-1, 2, 2)
    rh, zh = r.cpu().numpy(), z.cpu().numpy()
    y = _coloured_line_gauss_seidel(rowptr, col, blocks, lines, rh[:2 * nc].reshape(nc, 2), sweeps)
    err = np.abs(zh[:2 * nc] - y.ravel()).max() / np.abs(y).max()
    assert err <= 1e-11, f"{sweeps} sweeps: line stage differs from line Gauss-Seidel by {err:.2e}"
    # the well unknowns: each well's own block (equations x [qWs, qOs, bhp])
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, dt, cdp)
    Jd = J_ref.tocsr()
    for w in range(nw):
        idx = 2 * nc + np.arange(3) * nw + w
        zw = np.linalg.solve(Jd[idx][:, idx].toarray(), rh[2 * nc + 3 * w:2 * nc + 3 * w + 3])
        assert np.abs(zh[2 * nc + 3 * w:2 * nc + 3 * w + 3] - zw).max() <= 1e-9 * np.abs(zw).max(), w


def _line_stage_application(nx, ny, nz, lines, sweeps, seed=5, segments=False):
    """One application of the line local stage alone (FP64, no pressure
    stage, no graph replay) on the given lines, and the dense coloured line
    Gauss-Seidel restatement of it on the exported Jacobian blocks."""
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    inp = small_gravity_inputs(nx=nx, ny=ny, nz=nz)
    mo, so = build_oracle(inp)
    st = so.copy()
    rng = np.random.default_rng(seed)
    st.pressure = st.pressure + rng.normal(0, 2e5, mo.nc)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, mo.nc), 0.0, 1.0)
    dt = 10 * og.DAY
    cdp = mo.compute_cdp_per_well(so)
    solver = CPRSolver(rtol=1e-10, precision=64, local_sweeps=sweeps, use_pressure_stage=False, use_graphs=False,
                       line_segments=segments)
    md, _ = build_device(inp, solver)
    if lines is not None:
        ptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum([len(s) for s in lines])]), dtype=np.int32)
        cells = np.ascontiguousarray(np.concatenate(lines), dtype=np.int32)
        _lib.check(md.lib.pmx_ow_set_lines(md.h, len(lines), _lib.ptr(ptr), _lib.ptr(cells)), "pmx_ow_set_lines")
    else:
        lines = [np.array([i + nx * (j + ny * k) for k in range(nz)]) for j in range(ny) for i in range(nx)]
    dev = torch.device("cuda")
    md.assemble_device(DeviceOWState.from_host(so, dev), DeviceOWState.from_host(st, dev), dt,
                       torch.as_tensor(np.concatenate(cdp), device=dev))
    ok, pmean = md.converged(dt, DeviceOWState.from_host(st, dev))
    opts = solver.options()
    info = _lib.SolveInfo()
    rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(opts), pmean, _lib.ptr(md._dx), ctypes.byref(info))
    assert rc in (0, 1), rc
    nc, nw = mo.nc, len(mo.wells)
    r = torch.as_tensor(np.random.default_rng(9).normal(size=2 * nc + 3 * nw), device=dev)
    z = torch.empty_like(r)
    _lib.check(md.lib.pmx_ow_precondition(md.h, ctypes.byref(opts), _lib.ptr(r), _lib.ptr(z)),
               "pmx_ow_precondition")
    nnzb = ctypes.c_int32()
    _lib.check(md.lib.pmx_ow_nnzb(md.h, ctypes.byref(nnzb)))
    rowptr = np.zeros(nc + 1, np.int32)
    col = np.zeros(nnzb.value, np.int32)
    _lib.check(md.lib.pmx_ow_export_pattern(md.h, _lib.ptr(rowptr), _lib.ptr(col)))
    vals = torch.empty(nnzb.value * 4, dtype=torch.float64, device=dev)
    _lib.check(md.lib.pmx_ow_export_blocks(md.h, _lib.ptr(vals)))
    blocks = vals.cpu().numpy().reshape(-1, 2, 2)
    rh, zh = r.cpu().numpy(), z.cpu().numpy()
    y = _coloured_line_gauss_seidel(rowptr, col, blocks, lines, rh[:2 * nc].reshape(nc, 2), sweeps)
    return zh[:2 * nc], y.ravel()


@pytest.mark.parametrize("segments", [True, False])
@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("cut", [False, True])
def test_long_line_stage_matches_coloured_line_gauss_seidel(cut, sweeps, segments):
    """Columns of 85 cells, each solved by one twisted walk or (segments)
    by four segments cut at three separator cells, with spikes and a Schur
    system on the separators; the cut variant also makes lines of 1 to 85
    cells, so segmented lines of 32-63 cells (two segments) and whole lines
    share waves and colours.  One application equals the dense coloured
    block line Gauss-Seidel to 1e-11 (FP64): both are exact line solves."""
    nx, ny, nz = 6, 5, 85
    lines = _irregular_lines(nx, ny, nz, seed=4) if cut else None
    if cut:
        lengths = [len(s) for s in lines]
        assert max(lengths) >= 64 and any(32 <= n < 64 for n in lengths) and min(lengths) < 32
    print(f"long line stage: cut={cut} sweeps={sweeps} segments={segments}", flush=True)
    z, y = _line_stage_application(nx, ny, nz, lines, sweeps, segments=segments)
    err = np.abs(z - y).max() / np.abs(y).max()
    assert err <= 1e-11, f"cut={cut} sweeps={sweeps}: line stage differs from line Gauss-Seidel by {err:.2e}"


@pytest.mark.parametrize("prec", [64, 32])
def test_partitioned_lines_solve_matches_direct(prec):
    """CPR-FGMRES with the partitioned line stage on 85-cell columns solves a
    perturbed Newton system to the direct solution within 1e-7 relative
    (the reference's CPR contract, linear_solvers.py:639-649)."""
    from pymrst_amd import _lib
    from pymrst_amd.ow import CPRSolver, DeviceOWState
    nx, ny, nz = 6, 5, 85
    inp = small_gravity_inputs(nx=nx, ny=ny, nz=nz)
    mo, so = build_oracle(inp)
    st = so.copy()
    rng = np.random.default_rng(13)
    st.pressure = st.pressure + rng.normal(0, 2e5, mo.nc)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, mo.nc), 0.0, 1.0)
    dt = 10 * og.DAY
    cdp = mo.compute_cdp_per_well(so)
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, dt, cdp)
    x_ref = spla.spsolve(J_ref.tocsc(), -r_ref)
    perm = device_to_reference_perm(mo.nc, len(mo.wells))
    md, _ = build_device(inp, CPRSolver(rtol=1e-10, precision=prec, line_segments=True))
    dev = torch.device("cuda")
    md.assemble_device(DeviceOWState.from_host(so, dev), DeviceOWState.from_host(st, dev), dt,
                       torch.as_tensor(np.concatenate(cdp), device=dev))
    ok, pmean = md.converged(dt, DeviceOWState.from_host(st, dev))
    info = _lib.SolveInfo()
    print(f"partitioned lines CPR solve FP{prec}", flush=True)
    rc = md.lib.pmx_ow_solve(md.h, ctypes.byref(md.linear_solver.options()), pmean, _lib.ptr(md._dx),
                             ctypes.byref(info))
    assert rc == 0, f"FP{prec}: not converged ({info.iterations} its, rel {info.rel_residual:.1e})"
    x = md._dx.cpu().numpy()
    err = float(np.abs(x - x_ref[perm]).max() / np.abs(x_ref).max())
    assert err <= 1e-7, f"FP{prec}: CPR is not the direct solution: {err:.2e} ({info.iterations} its)"
