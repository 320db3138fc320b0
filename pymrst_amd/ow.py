"""Fully-implicit oil-water model on the MI355X: the reference's
TwoPhaseOilWaterModelAD call surface over the pmx HIP library.

Reference: pymrst/modules/ad_blackoil/models/two_phase_oil_water_model.py
(initial_state :172-186, _compute_cdp_per_well :226-268, assemble :493-546,
_converged :549-567, _solve_increment :569-603, step :605-657) and
pymrst/modules/ad_core/simulators/simulate_schedule_ad.py:17-80.

Residual, Jacobian, linear solve (CPR-FGMRES) and the limited Newton update
run on the GPU; this module keeps the host control flow of the Newton loop
(convergence decisions, time-step cuts) and the once-per-step well
hydrostatics.  States live on the device as torch tensors between calls.
"""
import ctypes
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .grid import BARSA, Grid, Rock, Well, cartesian_lines, compute_trans, pore_volume, tpfa_connections


@dataclass
class SimpleADIFluid:
    """init_simple_adi_fluid for phases 'WO'
    (pymrst/modules/ad_props/init_simple_adi_fluid.py:94-322):
    b_a = exp(c_a (p - p_ref)), kr_a = s_a^n_a, constant viscosity; ``c_r``
    adds the pore-volume multiplier 1 + c_r (p - p_ref) (:313-320)."""
    mu: Sequence[float]
    rho: Sequence[float]
    n: Sequence[float] = (1.0, 1.0)
    c: Sequence[float] = (0.0, 0.0)
    p_ref: float = 0.0
    c_r: Optional[float] = None

    @property
    def rhoS(self):
        return np.asarray(self.rho, float)

    def b(self, i, p):
        c = float(self.c[i])
        return np.ones_like(p) if c == 0.0 else np.exp(c * (np.asarray(p) - self.p_ref))

    def params(self) -> np.ndarray:
        return np.array([self.mu[0], self.mu[1], self.rho[0], self.rho[1], self.n[0], self.n[1],
                         self.c[0], self.c[1], self.p_ref, -1.0 if self.c_r is None else float(self.c_r)],
                        dtype=np.float64)


def init_simple_adi_fluid(mu, rho, n=(1.0, 1.0), c=(0.0, 0.0), p_ref=0.0, c_r=None) -> SimpleADIFluid:
    return SimpleADIFluid(mu=[float(v) for v in mu], rho=[float(v) for v in rho], n=[float(v) for v in n],
                          c=[float(v) for v in c], p_ref=float(p_ref), c_r=c_r)


@dataclass
class OWState:
    """Host state, as the reference's OWState (two_phase_oil_water_model.py:61-74)."""
    pressure: np.ndarray
    sW: np.ndarray
    qWs: np.ndarray
    qOs: np.ndarray
    bhp: np.ndarray

    def copy(self):
        return OWState(*(np.array(v, float, copy=True) for v in
                         (self.pressure, self.sW, self.qWs, self.qOs, self.bhp)))


@dataclass
class DeviceOWState:
    """The same state held in device tensors."""
    p: torch.Tensor
    sw: torch.Tensor
    qws: torch.Tensor
    qos: torch.Tensor
    bhp: torch.Tensor

    def copy(self):
        return DeviceOWState(*(t.clone() for t in (self.p, self.sw, self.qws, self.qos, self.bhp)))

    def to_host(self) -> OWState:
        return OWState(*(t.cpu().numpy() for t in (self.p, self.sw, self.qws, self.qos, self.bhp)))

    @staticmethod
    def from_host(st: OWState, device) -> "DeviceOWState":
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
        return DeviceOWState(f(st.pressure), f(st.sW), f(st.qWs), f(st.qOs), f(st.bhp))


@dataclass
class CPRSolver:
    """Options of the device CPR-FGMRES solve.  Defaults follow the
    reference's CPRSolver (linear_solvers.py:113-120): rtol 1e-8, restart 50,
    maxiter 200.  ``precision=32`` runs the preconditioner (pressure AMG,
    block-Jacobi, inner SpMV) in FP32 inside the FP64 flexible GMRES."""
    rtol: float = 1e-8
    maxiter: int = 200
    restart: int = 50
    precision: int = 64
    use_pressure_stage: bool = True
    pre_smooth: int = 2
    post_smooth: int = 2
    jacobi_omega: float = 1.0
    coarse_alpha: float = 1.5
    theta: float = 0.5
    coarse_size: int = 128
    pairwise_passes: int = 0
    reorthogonalize: bool = True
    verbose: int = 0
    local_stage: str = "line"
    pressure_cycles: int = 1
    local_sweeps: int = 4
    use_graphs: bool = True
    coarse_block_rows: int = 0
    line_segments: bool = False
    # launch shapes (AMG level layouts and folding, SpMV lanes per row) chosen
    # by timing at the first setup of each size class
    autotune: bool = False
    level_shape: int = -1      # 0-6: one launch shape for every AMG level (testing)
    dot_shape: int = 0         # 2, 4, 8, 104, 108: the projections' shape (testing)
    # a pressure-stage solver with solve_linear_system(A_p, r) -> x, called on
    # the equilibrated A_p = W J P in every preconditioner application
    # instead of the device AMG (linear_solvers.py:124-131, 253-256)
    elliptic_solver: object = None
    last_iterations: int = 0
    last_fallback: tuple = ()
    history: list = field(default_factory=list)

    def solve(self, jac, rhs, model=None, state=None, dt=None):
        """The reference's seam (linear_solvers.py:426-545): solve the
        caller's Jacobian (SciPy sparse, the reference's variable-blocked
        order) by CPR-FGMRES on the device; returns the increment as a NumPy
        array.  Non-convergence raises RuntimeError.  A model without .G,
        .n_phases and .fluid.{disgas, vapoil, b} gets the reference's loud
        fallback (a RuntimeWarning) to a direct solve, on the device; no
        model, the same solve without the warning."""
        from .linear import solve_reference_system
        return solve_reference_system(self, jac, rhs, model, state)

    def solve_linear_system(self, A, b):
        """A x = b without a model: the direct solve, on the device
        (linear_solver_ad.py:807)."""
        from .linear import device_direct_solve
        self.last_iterations = 0
        return device_direct_solve(A, b)

    def solve_linear_problem(self, problem, model=None):
        """The AD-OO seam (linear_solver_ad.py:832-893): (dx per primary
        variable, the increment, a report with MRST's fields)."""
        from .linear import solve_linear_problem
        return solve_linear_problem(self, problem, model)

    def get_description(self) -> str:
        return f"device CPR-FGMRES (FP{self.precision} preconditioner)"

    solveLinearSystem = solve_linear_system
    solveLinearProblem = solve_linear_problem

    def options(self) -> _lib.SolverOptions:
        o = _lib.solver_defaults()
        o.rtol, o.maxiter, o.restart, o.precision = self.rtol, self.maxiter, self.restart, self.precision
        o.use_pressure_stage = int(self.use_pressure_stage)
        o.pre_smooth, o.post_smooth = self.pre_smooth, self.post_smooth
        o.jacobi_omega, o.coarse_alpha, o.theta = self.jacobi_omega, self.coarse_alpha, self.theta
        o.coarse_size, o.pairwise_passes = self.coarse_size, self.pairwise_passes
        o.reorthogonalize, o.verbose = int(self.reorthogonalize), self.verbose
        o.local_stage = {"block_jacobi": 0, "line": 1, "ilu": 2}[self.local_stage]
        o.pressure_cycles = self.pressure_cycles
        o.local_sweeps = self.local_sweeps
        o.use_graphs = int(self.use_graphs)
        o.coarse_block_rows = self.coarse_block_rows
        o.line_segments = 1 if self.line_segments else 0
        o.autotune = 1 if self.autotune else 0
        o.level_shape = self.level_shape
        o.dot_shape = self.dot_shape
        return o


class TwoPhaseOilWaterModelAD:
    """Drop-in for the reference's TwoPhaseOilWaterModelAD on one GPU.

    Differences from the reference, all deliberate:
    * the linear solver is always the device CPR-FGMRES (``linear_solver``
      is a :class:`CPRSolver`; the reference's default is a direct solve);
    * a linear solve that does not converge makes ``step`` report failure,
      so the schedule driver cuts the time step (the legacy model would raise);
    * ``c_r`` on the fluid enables the pressure-dependent pore volume of the
      AD-OO oil-water equations.
    """

    n_phases = 2

    def __init__(self, g: Grid, rock: Rock, fluid: SimpleADIFluid, wells: Sequence[Well], gravity: float = 0.0,
                 ds_max: float = 0.2, dp_max_rel: float = 0.2, tol_cnv: float = 1e-3, tol_mb: float = 1e-7,
                 tol_well: float = 1e-6, max_iterations: int = 25, linear_solver: Optional[CPRSolver] = None,
                 device: Optional[torch.device] = None, T: Optional[np.ndarray] = None,
                 N: Optional[np.ndarray] = None):
        if not torch.cuda.is_available():
            raise _lib.PmxError("TwoPhaseOilWaterModelAD needs a GPU (no CPU fallback)")
        self.lib = _lib.load()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.g, self.rock, self.fluid = g, rock, fluid
        self.wells = list(wells)
        for w in self.wells:
            if w.type not in ("bhp", "rate"):
                raise ValueError(f"unsupported well control '{w.type}'")
        own_connections = N is None or T is None
        if own_connections:
            N, T, _, _ = tpfa_connections(g, compute_trans(g, rock))
        self.N = np.ascontiguousarray(N, dtype=np.int32)
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        self.pv = pore_volume(g, rock)
        self.pv_sum = float(self.pv.sum())
        self.gravity = float(gravity)
        z = g.cell_centroids[:, 2]
        self.gdz = (gravity * (z[self.N[:, 1]] - z[self.N[:, 0]]) if gravity
                    else np.zeros(self.N.shape[0]))
        self.ds_max, self.dp_max_rel = ds_max, dp_max_rel
        self.tol_cnv, self.tol_mb, self.tol_well = tol_cnv, tol_mb, tol_well
        self.max_iterations = max_iterations
        self.linear_solver = linear_solver if linear_solver is not None else CPRSolver()
        self.nc = g.num_cells
        self.nw = len(self.wells)
        self._last_cdp = None
        self.stats = {"newton": 0, "linear": 0, "assemble_s": 0.0, "solve_s": 0.0}
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
            h = ctypes.c_void_p()
            fp = fluid.params()
            _lib.check(self.lib.pmx_ow_create(self.device.index or 0, ctypes.c_void_p(self.stream.cuda_stream),
                                              _lib.ptr(fp), ctypes.byref(h)), "pmx_ow_create")
            self.h = h
            pv = np.ascontiguousarray(self.pv, dtype=np.float64)
            gdz = np.ascontiguousarray(self.gdz, dtype=np.float64)
            _lib.check(self.lib.pmx_ow_set_topology(h, self.nc, self.N.shape[0], _lib.ptr(self.N),
                                                    _lib.ptr(self.T), _lib.ptr(gdz), _lib.ptr(pv)),
                       "pmx_ow_set_topology")
            self._set_wells()
            if own_connections and getattr(g, "cartDims", None) is not None:
                axis, ptr, cells = cartesian_lines(g, self.N, self.T)
                _lib.check(self.lib.pmx_ow_set_lines(h, ptr.size - 1, _lib.ptr(ptr), _lib.ptr(cells)),
                           "pmx_ow_set_lines")
                self.line_axis = axis
        n = 2 * self.nc + 3 * self.nw
        self._dx = torch.zeros(n, dtype=torch.float64, device=self.device)

    def _set_wells(self):
        nw = self.nw
        counts = [w.cells.size for w in self.wells]
        self.perf_ptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum(counts)]), dtype=np.int32)
        self.perf_cells = (np.ascontiguousarray(np.concatenate([w.cells for w in self.wells]), dtype=np.int32)
                           if nw else np.zeros(0, np.int32))
        wi = (np.ascontiguousarray(np.concatenate([w.WI for w in self.wells]), dtype=np.float64)
              if nw else np.zeros(0))
        types = np.array([0 if w.type == "bhp" else 1 for w in self.wells], dtype=np.int32)
        vals = np.array([w.val for w in self.wells], dtype=np.float64)
        compi = np.ascontiguousarray(np.array([[w.compi[0], w.compi[1]] for w in self.wells],
                                              dtype=np.float64).reshape(-1))
        _lib.check(self.lib.pmx_ow_set_wells(self.h, nw, _lib.ptr(self.perf_ptr), _lib.ptr(self.perf_cells),
                                             _lib.ptr(wi), _lib.ptr(types), _lib.ptr(vals), _lib.ptr(compi)),
                   "pmx_ow_set_wells")
        self.perf_cells_t = torch.as_tensor(self.perf_cells.astype(np.int64), device=self.device)
        self._cdp_dev = torch.zeros(max(int(self.perf_ptr[-1]), 1), dtype=torch.float64, device=self.device)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and getattr(self, "lib", None) is not None:
            try:
                self.lib.pmx_ow_destroy(h)
            except Exception:
                pass
            self.h = None

    def reset(self):
        """Forget per-run state (the previous step's well cdp) so a new run
        from an initial state behaves like a fresh model."""
        self._last_cdp = None
        self._cdp_dev.zero_()
        self.stats = {"newton": 0, "linear": 0, "assemble_s": 0.0, "solve_s": 0.0}

    def run_memory(self) -> dict:
        """What a run carries from one step to the next besides the state:
        the previous step's per-connection well cdp, which seeds the next
        step's connection rates (the reference keeps it on the model,
        two_phase_oil_water_model.py:254-283).  A restart that restores it
        continues the run exactly."""
        return {} if self._last_cdp is None else {"well_cdp": [np.array(c, float) for c in self._last_cdp]}

    def restore_run_memory(self, memory: Optional[dict]):
        cdp = (memory or {}).get("well_cdp")
        if cdp is None:
            self._last_cdp = None
            self._cdp_dev.zero_()
            return
        if len(cdp) != self.nw or any(np.size(c) != w.cells.size for c, w in zip(cdp, self.wells)):
            raise ValueError("stored well cdp does not match this model's wells")
        self._last_cdp = [np.asarray(c, float) for c in cdp]
        flat = np.concatenate(self._last_cdp)
        self._cdp_dev[:flat.size].copy_(torch.as_tensor(flat, dtype=torch.float64))

    def profile(self, enable: bool):
        _lib.check(self.lib.pmx_ow_profile(self.h, int(enable)), "pmx_ow_profile")

    def launch_shapes(self) -> dict:
        return _lib.read_launch_shapes(self.lib.pmx_ow_launch_shapes, self.h)

    def profile_read(self) -> dict:
        out = np.zeros(4)
        _lib.check(self.lib.pmx_ow_profile_read(self.h, _lib.ptr(out)), "pmx_ow_profile_read")
        return {"spmv_count": int(out[0]), "spmv_ms": float(out[1]), "prec_count": int(out[2]),
                "prec_ms": float(out[3])}

    def time_line_solve(self, applications: int = 20) -> dict:
        """Time the line Gauss-Seidel solve kernel launch by launch (HIP
        events on the model's stream) over `applications` preconditioner
        applications without graph replay, on the setup of the last linear
        solve; returns the launches timed, their total ms and their
        algorithmic bytes."""
        o = self.linear_solver.options()
        o.use_graphs = 0
        n = 2 * self.nc + 3 * self.nw
        with torch.cuda.device(self.device):
            gen = torch.Generator(device=self.device).manual_seed(0)
            r = torch.randn(n, dtype=torch.float64, device=self.device, generator=gen)
            z = torch.empty_like(r)
            _lib.check(self.lib.pmx_ow_profile(self.h, 1), "pmx_ow_profile")
            out = np.zeros(3)
            try:
                for _ in range(applications):
                    _lib.check(self.lib.pmx_ow_precondition(self.h, ctypes.byref(o), _lib.ptr(r), _lib.ptr(z)),
                               "pmx_ow_precondition")
                _lib.check(self.lib.pmx_ow_profile_line_gs(self.h, _lib.ptr(out)), "pmx_ow_profile_line_gs")
            finally:
                self.lib.pmx_ow_profile(self.h, 0)
        return {"launches": int(out[0]), "ms": float(out[1]), "bytes": float(out[2])}

    @property
    def nnzb(self) -> int:
        v = ctypes.c_int32()
        _lib.check(self.lib.pmx_ow_nnzb(self.h, ctypes.byref(v)), "pmx_ow_nnzb")
        return int(v.value)

    # -- reference attribute contract ------------------------------------
    @property
    def G(self):
        return self.g

    def initial_state(self, p0: float, sw0: float) -> OWState:
        """initial_state (:172-186)."""
        bhp = np.array([w.val if w.type == "bhp" else p0 for w in self.wells], float)
        qws = np.array([w.val * w.compi[0] if w.type == "rate" else 0.0 for w in self.wells], float)
        qos = np.array([w.val * w.compi[1] if w.type == "rate" else 0.0 for w in self.wells], float)
        return OWState(np.full(self.nc, float(p0)), np.full(self.nc, float(sw0)), qws, qos, bhp)

    def device_state(self, st) -> DeviceOWState:
        return st if isinstance(st, DeviceOWState) else DeviceOWState.from_host(st, self.device)

    # -- device primitives -------------------------------------------------
    def assemble_device(self, st0: DeviceOWState, st: DeviceOWState, dt: float, cdp: Optional[torch.Tensor]):
        P = _lib.ptr
        _lib.check(self.lib.pmx_ow_assemble(self.h, P(st.p), P(st.sw), P(st.qws), P(st.qos), P(st.bhp),
                                            P(st0.p), P(st0.sw), float(dt), P(cdp)), "pmx_ow_assemble")

    def residual_device(self) -> torch.Tensor:
        r = torch.empty(2 * self.nc + 3 * self.nw, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.pmx_ow_residual(self.h, _lib.ptr(r)), "pmx_ow_residual")
        return r

    def convergence_values(self, st: DeviceOWState):
        out7 = np.zeros(7)
        wells = np.zeros(max(3 * self.nw, 1))
        _lib.check(self.lib.pmx_ow_convergence(self.h, _lib.ptr(st.p), _lib.ptr(out7), _lib.ptr(wells)),
                   "pmx_ow_convergence")
        return out7, wells[:3 * self.nw]

    def converged(self, dt: float, st: DeviceOWState):
        """_converged (:549-567) from device reductions; returns (ok, pmean)."""
        out7, wr = self.convergence_values(st)
        nc = self.nc
        bw, bo = out7[4] / nc, out7[5] / nc
        cnv = max(bw * dt * out7[0], bo * dt * out7[1])
        mb = max(bw * dt * abs(out7[2]) / self.pv_sum, bo * dt * abs(out7[3]) / self.pv_sum)
        ok = bool(cnv <= self.tol_cnv and mb <= self.tol_mb)
        if self.nw:
            rate = [abs(w.val) for w in self.wells if w.type == "rate"]
            q_ref = max(max(rate), 1e-12) if rate else 1.0
            eq_w, eq_o, ctrl = wr[0::3], wr[1::3], wr[2::3]
            well_ok = (np.abs(eq_w).max() <= self.tol_well * q_ref and np.abs(eq_o).max() <= self.tol_well * q_ref)
            scale = np.array([BARSA if w.type == "bhp" else q_ref for w in self.wells])
            ok = ok and bool(well_ok) and bool(np.all(np.abs(ctrl) <= self.tol_well * scale))
        self.last_cnv_mb = (cnv, mb)
        return ok, out7[6] / nc

    def solve_and_update(self, st: DeviceOWState, pmean: float, scale: float = 1.0) -> bool:
        info = _lib.SolveInfo()
        opts = self.linear_solver.options()
        t0 = time.perf_counter()
        rc = self.lib.pmx_ow_solve(self.h, ctypes.byref(opts), float(pmean), _lib.ptr(self._dx),
                                   ctypes.byref(info))
        self.stats["solve_s"] += time.perf_counter() - t0
        _lib.check(rc, "pmx_ow_solve")
        self.linear_solver.last_iterations = int(info.iterations)
        self.linear_solver.history.append((int(info.iterations), float(info.rel_residual), float(info.setup_ms),
                                           float(info.solve_ms)))
        self.stats["linear"] += int(info.iterations)
        if rc != 0:
            return False
        self.apply_increment(st, scale)
        return True

    def apply_increment(self, st: DeviceOWState, scale: float = 1.0):
        """st += scale * (the last Newton increment), with the update limits."""
        P = _lib.ptr
        _lib.check(self.lib.pmx_ow_update(self.h, P(self._dx), float(scale), self.ds_max, self.dp_max_rel, P(st.p),
                                          P(st.sw), P(st.qws), P(st.qos), P(st.bhp)), "pmx_ow_update")

    # -- wells: hydrostatic connection pressure drops, once per step ----------
    def compute_cdp_per_well(self, st0: DeviceOWState) -> Optional[torch.Tensor]:
        """_compute_cdp_per_well (:226-268): per-connection mixture density
        from the connection rates at st0 (evaluated with the previous cdp),
        cumulated down each well from its reference depth."""
        if not self.gravity or not self.nw:
            return None
        prev = self._cdp_dev if self._last_cdp is not None else None
        self.assemble_device(st0, st0, 1.0, prev)
        rates = np.zeros(2 * int(self.perf_ptr[-1]))
        _lib.check(self.lib.pmx_ow_connection_rates(self.h, _lib.ptr(rates)), "pmx_ow_connection_rates")
        cq_w, cq_o = rates[0::2], rates[1::2]
        p_perf = st0.p[self.perf_cells_t].cpu().numpy()
        bw, bo = self.fluid.b(0, p_perf), self.fluid.b(1, p_perf)
        rho = self.fluid.rhoS
        out = []
        for k, w in enumerate(self.wells):
            a, b = self.perf_ptr[k], self.perf_ptr[k + 1]
            sw_ = np.cumsum(cq_w[a:b][::-1])[::-1]
            so_ = np.cumsum(cq_o[a:b][::-1])[::-1]
            tot = sw_ + so_
            dead = tot == 0.0
            safe = np.where(dead, 1.0, tot)
            mw = np.where(dead, w.compi[0], sw_ / safe)
            mo = np.where(dead, w.compi[1], so_ / safe)
            rho_mix = (mw * rho[0] + mo * rho[1]) / (mw / bw[a:b] + mo / bo[a:b])
            dz = np.diff(np.concatenate(([0.0], w.dZ)))
            out.append(np.cumsum(self.gravity * rho_mix * dz))
        self._last_cdp = out
        cdp = np.concatenate(out)
        self._cdp_dev[:cdp.size].copy_(torch.as_tensor(cdp, dtype=torch.float64))
        return self._cdp_dev

    # -- ministep protocol of pymrst_amd.nonlinear.NonLinearSolver --------------
    def begin_ministep(self, state0: DeviceOWState):
        self._ms_cdp = self.compute_cdp_per_well(state0)

    def linearize(self, state0: DeviceOWState, st: DeviceOWState, dt: float):
        """Assemble at st and test convergence: (converged, {CNV, MB})."""
        self.assemble_device(state0, st, dt, self._ms_cdp)
        ok, self._ms_pmean = self.converged(dt, st)
        cnv, mb = self.last_cnv_mb
        return ok, {"CNV": float(cnv), "MB": float(mb)}

    def newton_update(self, st: DeviceOWState, scale: float = 1.0) -> bool:
        self.stats["newton"] += 1
        return self.solve_and_update(st, self._ms_pmean, scale)

    @property
    def last_linear_solve(self):
        h = self.linear_solver.history[-1]
        return int(h[0]), float(h[1])

    # -- Newton ----------------------------------------------------------------
    def step_device(self, state: DeviceOWState, dt: float):
        """step (:605-657) on device state: returns (state, converged, its)."""
        st0 = state
        st = state.copy()
        cdp = self.compute_cdp_per_well(st0)
        for it in range(self.max_iterations):
            t0 = time.perf_counter()
            self.assemble_device(st0, st, dt, cdp)
            ok, pmean = self.converged(dt, st)
            self.stats["assemble_s"] += time.perf_counter() - t0
            if ok:
                return st, True, it
            self.stats["newton"] += 1
            if not self.solve_and_update(st, pmean):
                return st, False, it + 1
        self.assemble_device(st0, st, dt, cdp)
        ok, _ = self.converged(dt, st)
        return st, ok, self.max_iterations

    def step(self, state, dt: float):
        if isinstance(state, DeviceOWState):
            return self.step_device(state, dt)
        st, ok, its = self.step_device(DeviceOWState.from_host(state, self.device), dt)
        return st.to_host(), ok, its


def simulate_schedule_ad(state0, model, timesteps: Sequence[float], max_cuts: int = 6,
                         verbose: bool = False, observer=None, keep_device: bool = False):
    """simulate_schedule_ad (simulate_schedule_ad.py:17-80): report steps,
    halving cuts on Newton failure; returns (states, iterations per step)."""
    def _observe(event, **payload):
        if observer is None:
            return
        try:
            observer(event, payload)
        except Exception:
            pass

    state = model.device_state(state0)
    states: List = []
    its_per_step = np.zeros(len(timesteps), dtype=np.int64)
    for i, dt_report in enumerate(timesteps):
        _observe("report_step_started", report_step=int(i), dt=float(dt_report))
        t_local, its = 0.0, 0
        dt, cuts = float(dt_report), 0
        while t_local < dt_report - 1e-12 * dt_report:
            dt = min(dt, dt_report - t_local)
            _observe("ministep_started", report_step=int(i), dt=float(dt))
            cand, ok, nit = model.step_device(state, dt)
            its += nit
            _observe("step_accepted" if ok else "step_rejected", report_step=int(i), dt=float(dt),
                     iterations=int(nit))
            if ok:
                state = cand
                t_local += dt
                dt = min(2.0 * dt, dt_report - t_local) if cuts else dt
            else:
                cuts += 1
                if cuts > max_cuts:
                    raise RuntimeError(f"step {i}: Newton failed after {max_cuts} timestep cuts")
                dt /= 2.0
                _observe("timestep_cut", report_step=int(i), cut=int(cuts), dt=float(dt))
                if verbose:
                    print(f"  step {i}: cutting dt to {dt:.3e}")
        its_per_step[i] = its
        states.append(state if keep_device else state.to_host())
        if verbose:
            print(f"step {i + 1}/{len(timesteps)}: {its} iterations")
    return states, its_per_step


simulateScheduleAD = simulate_schedule_ad
