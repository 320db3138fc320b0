"""ctypes binding of the pmx C ABI (include/pmx.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both carry the soname libamdhip64.so.7).  A missing library
is an error: there is no CPU fallback for the device path.
"""
import ctypes
from pathlib import Path

import numpy as np

import torch  # noqa: F401  (loads the HIP runtime the library links against)

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_pmx.so"


class SolverOptions(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double),
        ("restart", ctypes.c_int32),
        ("maxiter", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("use_pressure_stage", ctypes.c_int32),
        ("pre_smooth", ctypes.c_int32),
        ("post_smooth", ctypes.c_int32),
        ("jacobi_omega", ctypes.c_double),
        ("coarse_alpha", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("coarse_size", ctypes.c_int32),
        ("pairwise_passes", ctypes.c_int32),
        ("reorthogonalize", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("local_stage", ctypes.c_int32),
        ("pressure_cycles", ctypes.c_int32),
        ("local_sweeps", ctypes.c_int32),
        ("use_graphs", ctypes.c_int32),
        ("coarse_block_rows", ctypes.c_int32),
        ("line_segments", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("level_shape", ctypes.c_int32),
        ("dot_shape", ctypes.c_int32),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("restarts", ctypes.c_int32),
        ("rel_residual", ctypes.c_double),
        ("setup_ms", ctypes.c_double),
        ("solve_ms", ctypes.c_double),
        ("levels", ctypes.c_int32),
        ("coarse_n", ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double

#: name -> (restype, argtypes); every symbol declared in include/pmx.h
SIGNATURES = {
    "pmx_last_error": (ctypes.c_char_p, []),
    "pmx_solver_defaults": (None, [ctypes.POINTER(SolverOptions)]),
    "pmx_ow_create": (ctypes.c_int, [_I, _P, _P, ctypes.POINTER(_P)]),
    "pmx_ow_destroy": (ctypes.c_int, [_P]),
    "pmx_ow_set_topology": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P]),
    "pmx_ow_set_wells": (ctypes.c_int, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "pmx_ow_assemble": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _P]),
    "pmx_ow_residual": (ctypes.c_int, [_P, _P]),
    "pmx_ow_convergence": (ctypes.c_int, [_P, _P, _P, _P]),
    "pmx_ow_convergence_noalloc": (ctypes.c_int, [_P, _P, _P, _P]),
    "pmx_ow_set_controls": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "pmx_ow_convergence_current_pv": (ctypes.c_int, [_P, _P, _P, _P]),
    "pmx_ow_perf_volume_rates": (ctypes.c_int, [_P, _P]),
    "pmx_ow_solve": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), _D, _P, ctypes.POINTER(SolveInfo)]),
    "pmx_ow_update": (ctypes.c_int, [_P, _P, _D, _D, _D, _P, _P, _P, _P, _P]),
    "pmx_ow_matvec": (ctypes.c_int, [_P, _P, _P]),
    "pmx_ow_nnzb": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "pmx_ow_export_pattern": (ctypes.c_int, [_P, _P, _P]),
    "pmx_ow_export_blocks": (ctypes.c_int, [_P, _P]),
    "pmx_ow_connection_rates": (ctypes.c_int, [_P, _P]),
    "pmx_ow_set_lines": (ctypes.c_int, [_P, _I, _P, _P]),
    "pmx_ow_profile": (ctypes.c_int, [_P, _I]),
    "pmx_ow_profile_read": (ctypes.c_int, [_P, _P]),
    "pmx_ow_profile_line_gs": (ctypes.c_int, [_P, _P]),
    "pmx_ow_launch_shapes": (ctypes.c_int, [_P, _P, _I, _P]),
    "pmx_autotune_entries": (ctypes.c_int, [_P]),
    "pmx_autotune_reset": (ctypes.c_int, []),
    "pmx_autotune_save": (ctypes.c_int, [ctypes.c_char_p, _P]),
    "pmx_autotune_load": (ctypes.c_int, [ctypes.c_char_p, _P]),
    "pmx_ow_precondition": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), _P, _P]),
    "pmx_sys_create": (ctypes.c_int, [_I, _P, _I, _I, _P, _P, _I, _P, _P, ctypes.POINTER(_P)]),
    "pmx_sys_destroy": (ctypes.c_int, [_P]),
    "pmx_sys_set_values": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "pmx_sys_matvec": (ctypes.c_int, [_P, _P, _P]),
    "pmx_sys_solve": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), _P, _P, _P, ctypes.POINTER(SolveInfo)]),
    "pmx_ow_set_trans_fractions": (ctypes.c_int, [_P, _P]),
    "pmx_ow_set_segments": (ctypes.c_int, [_P, _I, _P, _P]),
    "pmx_ow_mask_residual": (ctypes.c_int, [_P, _P]),
    "pmx_ow_set_cell_dt": (ctypes.c_int, [_P, _P]),
    "pmx_ow_set_members": (ctypes.c_int, [_P, _I, _P]),
    "pmx_ow_member_convergence": (ctypes.c_int, [_P, _P, _P, _P]),
    "pmx_ow_adjoint_solve": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), ctypes.c_double, _P, _P,
                                            ctypes.POINTER(SolveInfo)]),
    "pmx_ow_prev_state_transpose": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, _P, _P]),
    "pmx_ow_dlogk": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pmx_bo_create": (ctypes.c_int, [_I, _P, _P, _I, _P, _I, _P, _P, _P, _I, _P, _I, _P, ctypes.POINTER(_P)]),
    "pmx_bo_destroy": (ctypes.c_int, [_P]),
    "pmx_bo_set_topology": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P]),
    "pmx_bo_set_wells": (ctypes.c_int, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "pmx_bo_set_lines": (ctypes.c_int, [_P, _I, _P, _P]),
    "pmx_bo_begin_step": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "pmx_bo_assemble": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _D]),
    "pmx_bo_residual": (ctypes.c_int, [_P, _P]),
    "pmx_bo_convergence": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "pmx_bo_solve": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), _D, _D, _P, ctypes.POINTER(SolveInfo)]),
    "pmx_bo_update": (ctypes.c_int, [_P, _P, _D, _P, _P, _P, _P, _P, _P]),
    "pmx_bo_set_controls": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "pmx_bo_set_connection_drops": (ctypes.c_int, [_P, _P]),
    "pmx_bo_perf_volume_rates": (ctypes.c_int, [_P, _P]),
    "pmx_bo_set_update_rule": (ctypes.c_int, [_P, _I, _D, _D]),
    "pmx_bo_cell_densities": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "pmx_bo_matvec": (ctypes.c_int, [_P, _P, _P]),
    "pmx_bo_nnzb": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "pmx_bo_export_pattern": (ctypes.c_int, [_P, _P, _P]),
    "pmx_bo_export_blocks": (ctypes.c_int, [_P, _P]),
    "pmx_bo_profile": (ctypes.c_int, [_P, _I]),
    "pmx_bo_profile_read": (ctypes.c_int, [_P, _P]),
    "pmx_bo_launch_shapes": (ctypes.c_int, [_P, _P, _I, _P]),
    "pmx_cp_process": (ctypes.c_int, [_I, _I, _I, _P, _P, _P, ctypes.POINTER(_P)]),
    "pmx_cp_destroy": (ctypes.c_int, [_P]),
    "pmx_cp_sizes": (ctypes.c_int, [_P, _P]),
    "pmx_cp_topology": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pmx_cp_geometry": (ctypes.c_int, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "pmx_incomp_create": (ctypes.c_int, [_I, _P, ctypes.POINTER(_P)]),
    "pmx_incomp_destroy": (ctypes.c_int, [_P]),
    "pmx_incomp_set_grid": (ctypes.c_int, [_P, _I, _I, _P, _P, _P, _P, _P]),
    "pmx_compute_trans": (ctypes.c_int, [_P, _P, _I, _P, _P]),
    "pmx_incomp_solve": (ctypes.c_int, [_P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P,
                                        ctypes.POINTER(SolverOptions), _P, _P, _P, _P, _P,
                                        ctypes.POINTER(SolveInfo)]),
    "pmx_incomp_assemble": (ctypes.c_int, [_P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _I, _P, _P, _I, _P, _P,
                                           _P, _P, _P, _P]),
    "pmx_incomp_nnz": (ctypes.c_int, [_P, ctypes.POINTER(_I)]),
    "pmx_sys_set_pressure_solver": (ctypes.c_int, [_P, _P, _P]),
    "pmx_incomp_export": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "pmx_incomp_solve_assembled": (ctypes.c_int, [_P, ctypes.POINTER(SolverOptions), _P, ctypes.POINTER(SolveInfo)]),
    "pmx_incomp_reconstruct": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


class PmxError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise PmxError(f"pmx HIP library not built: {LIB_PATH} missing "
                       "(run `python -m pymrst_amd.build`)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "pmx call") -> int:
    """Raise on a negative status; return 0/1 (converged / not converged)."""
    if rc < 0:
        msg = load().pmx_last_error().decode(errors="replace")
        raise PmxError(f"{what} failed: {msg}")
    return rc


def ptr(t) -> ctypes.c_void_p:
    """Device or host pointer of a tensor / numpy array (None -> NULL)."""
    if t is None:
        return ctypes.c_void_p(None)
    if isinstance(t, torch.Tensor):
        assert t.is_contiguous(), "tensor must be contiguous"
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


def solver_defaults() -> SolverOptions:
    o = SolverOptions()
    load().pmx_solver_defaults(ctypes.byref(o))
    return o


def read_launch_shapes(fn, h) -> dict:
    """The launch shapes a model's solver uses (pmx_*_launch_shapes): the
    SpMV's lanes per block row and, per AMG level above the coarse cycle,
    its rows, sweep layout, lanes per CSR row and folding."""
    out = np.zeros(512, np.int32)
    n = ctypes.c_int32()
    check(fn(h, ptr(out), out.size, ctypes.byref(n)), "launch_shapes")
    out = out[:n.value]
    nl = int(out[1])
    levels = out[2:2 + 5 * nl].reshape(-1, 5)
    return {"spmv_lanes": int(out[0]), "projection_elems": int(out[2 + 5 * nl]) if out.size > 2 + 5 * nl else 0,
            "levels": [{"rows": int(r), "layout": "ell" if e else "csr", "lanes": int(l), "folded": bool(f),
                        "tuned": bool(t)} for r, e, l, f, t in levels]}


def autotune_save(path) -> int:
    """Write the process-wide launch-shape table (pmx_autotune_save); returns
    the number of entries."""
    n = ctypes.c_int32()
    check(load().pmx_autotune_save(str(path).encode(), ctypes.byref(n)), "pmx_autotune_save")
    return n.value


def autotune_load(path) -> int:
    """Read a table written by autotune_save into the process-wide table
    (pmx_autotune_load): later setups of those size classes use its shapes
    without timing; returns the number of entries read."""
    n = ctypes.c_int32()
    check(load().pmx_autotune_load(str(path).encode(), ctypes.byref(n)), "pmx_autotune_load")
    return n.value
