"""Discrete adjoint of the device oil-water Newton run, for gradients of a
schedule objective with respect to every cell's log-permeability and to the
well controls.

Reference: pymrst/modules/optimization/adjoint.py.  The backward sweep is the
reference's (:163-225):

    (dR_n/dx_n)^T lam_n = -dJ_n/dx_n - (dR_{n+1}/dx_n)^T lam_{n+1}
    dJ/du_n             = (dR_n/du_n)^T lam_n      (-1 in the control rows, :107-121)

with dR_n/dx_{n-1} taken from the accumulation term only (:85-107).  The
reference's parameters are the well controls; the log-permeability gradient
adds dJ/dln k = sum_n (dR_n/dln k)^T lam_n, through the face
transmissibilities (T_h proportional to k, dT_f/dln k_c = T_f^2/T_h,c) and
the Peaceman well indices (proportional to k).  Everything per step runs on
the device: assembly at the stored states, the transposed CPR-FGMRES solve,
the previous-state coupling and the dR/dln k kernels.

States are kept at checkpoints (every ``checkpoint_every`` accepted
ministeps); the backward sweep recomputes the states in between by
re-running those ministeps, which reproduces them bit for bit (the device
Newton is deterministic).  The well head ``cdp`` of each ministep is stored
and held fixed, as in the forward run (the reference's adjoint assembles
with cdp = 0, adjoint.py:73-80); its own dependence on the previous state is
not differentiated, so gradients are exact only without gravity in the
wells.
"""
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import ctypes
import numpy as np
import torch

from . import _lib
from .grid import Grid, Rock, compute_trans, make_rock, tpfa_connections
from .ow import CPRSolver, DeviceOWState, TwoPhaseOilWaterModelAD


def transmissibility_fractions(g: Grid, t_half: np.ndarray, N: np.ndarray) -> np.ndarray:
    """Per internal connection (N0, N1): T_f / T_h at each side, the
    log-permeability derivative of T_f divided by T_f."""
    nf = g.num_faces
    th = np.zeros((nf, 2))
    faces = g.cell_faces.ravel()
    cells = np.repeat(np.arange(g.num_cells), g.cell_faces.shape[1])
    side = (g.neighbors[faces, 1] == cells).astype(np.int64)
    th[faces, side] = t_half
    internal = np.all(g.neighbors >= 0, axis=1)
    th = th[internal]
    Tf = 1.0 / (1.0 / th[:, 0] + 1.0 / th[:, 1])
    frac = np.stack([Tf / th[:, 0], Tf / th[:, 1]], axis=1)
    # the connection list of the model is the internal faces in face order
    assert frac.shape[0] == N.shape[0]
    return np.ascontiguousarray(frac.ravel())


@dataclass
class Ministep:
    dt: float
    report: int            # report step this ministep belongs to
    ends_report: bool      # the objective scores the state at its end
    cdp: Optional[torch.Tensor]


@dataclass
class ForwardRun:
    ministeps: List[Ministep]
    checkpoints: dict                  # ministep index -> state at its start (device)
    final: DeviceOWState
    report_states: List[DeviceOWState] = field(default_factory=list)
    newton: List[int] = field(default_factory=list)


class AdjointProblem:
    """One oil-water schedule on a fixed grid, wells and fluid, as a function
    of the log-permeability field (isotropy ratios fixed by ``aniso``).

    ``objective(n, qws, qos, bhp)`` returns J_n from the well state at the
    end of report step n as a torch scalar (differentiable in those
    tensors); ``objective_wells`` may be None to use the water-rate mismatch
    of the producers against ``targets[n, w]``.
    """

    def __init__(self, g: Grid, poro, fluid, well_specs: Sequence[dict], dts: Sequence[float], p0: float,
                 sw0: float, aniso=(1.0, 1.0, 0.1), gravity: float = 0.0, solver: Optional[CPRSolver] = None,
                 objective: Optional[Callable] = None, targets: Optional[np.ndarray] = None,
                 extra_newton: int = 2, checkpoint_every: int = 1, device=None):
        from .grid import add_well
        self.g, self.poro, self.fluid = g, np.asarray(poro, float), fluid
        self.well_specs = list(well_specs)
        self.dts = np.asarray(dts, float)
        self.p0, self.sw0 = float(p0), float(sw0)
        self.aniso = np.asarray(aniso, float)
        self.gravity = float(gravity)
        self.solver = solver if solver is not None else CPRSolver(rtol=1e-12, precision=64)
        self.extra_newton = int(extra_newton)
        self.checkpoint_every = max(1, int(checkpoint_every))
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self._add_well = add_well
        if objective is None:
            if targets is None:
                raise ValueError("give an objective or the producer water-rate targets")
            self.targets = torch.as_tensor(np.asarray(targets, float), device=self.device)
            prod = [k for k, s in enumerate(self.well_specs) if s.get("type", "bhp") == "bhp"]
            self.producers = torch.as_tensor(prod, dtype=torch.long, device=self.device)
            objective = self._rate_mismatch
        self.objective = objective
        self.model: Optional[TwoPhaseOilWaterModelAD] = None

    # -- objective -------------------------------------------------------------
    def _rate_mismatch(self, n, qws, qos, bhp):
        """sum_w (qWs_w - target_n,w)^2 over the producers."""
        d = qws[self.producers] - self.targets[n]
        return (d * d).sum()

    # -- model for a log-permeability field ----------------------------------------
    def build(self, logk: np.ndarray) -> TwoPhaseOilWaterModelAD:
        g = self.g
        perm = np.exp(np.asarray(logk, float))[:, None] * self.aniso[None, :]
        rock = make_rock(g, perm, self.poro)
        t_half = compute_trans(g, rock)
        N, T, _, _ = tpfa_connections(g, t_half)
        wells = None
        for s in self.well_specs:
            kind = s.get("type", "bhp")
            wells = self._add_well(wells, g, rock, s["cells"], kind, s["val"], radius=s.get("radius", 0.1),
                                   compi=np.asarray(s.get("compi", [1.0, 0.0]), float),
                                   sign=s.get("sign", 1 if kind == "rate" else -1), gravity=None)
        m = TwoPhaseOilWaterModelAD(g, rock, self.fluid, wells, gravity=self.gravity, linear_solver=self.solver,
                                    device=self.device, max_iterations=40)
        frac = transmissibility_fractions(g, t_half, m.N)
        _lib.check(m.lib.pmx_ow_set_trans_fractions(m.h, _lib.ptr(frac)), "pmx_ow_set_trans_fractions")
        self.model = m
        return m

    # -- forward ---------------------------------------------------------------
    def _newton(self, m: TwoPhaseOilWaterModelAD, st0: DeviceOWState, dt: float, cdp):
        """The model's Newton step (ow.py step_device) with a given cdp, then
        `extra_newton` further iterations so the state is converged to
        round-off (finite differences of J need that)."""
        st = st0.copy()
        for it in range(m.max_iterations):
            m.assemble_device(st0, st, dt, cdp)
            ok, pmean = m.converged(dt, st)
            if ok:
                for _ in range(self.extra_newton):
                    if not m.solve_and_update(st, pmean):
                        break
                    m.assemble_device(st0, st, dt, cdp)
                    _, pmean = m.converged(dt, st)
                return st, True, it
            if not m.solve_and_update(st, pmean):
                return st, False, it + 1
        return st, False, m.max_iterations

    def forward(self, logk: np.ndarray, max_cuts: int = 6) -> ForwardRun:
        m = self.build(logk)
        state = DeviceOWState.from_host(m.initial_state(self.p0, self.sw0), self.device)
        run = ForwardRun(ministeps=[], checkpoints={}, final=state)
        for n, dt_report in enumerate(self.dts):
            t_local, dt, cuts, its = 0.0, float(dt_report), 0, 0
            while t_local < dt_report - 1e-12 * dt_report:
                dt = min(dt, dt_report - t_local)
                cdp = m.compute_cdp_per_well(state)
                cdp = cdp.clone() if cdp is not None else None
                cand, ok, nit = self._newton(m, state, dt, cdp)
                its += nit
                if not ok:
                    cuts += 1
                    if cuts > max_cuts:
                        raise RuntimeError(f"forward step {n} did not converge")
                    dt /= 2.0
                    continue
                k = len(run.ministeps)
                if k % self.checkpoint_every == 0:
                    run.checkpoints[k] = state
                t_local += dt
                run.ministeps.append(Ministep(dt=dt, report=n, ends_report=t_local >= dt_report * (1 - 1e-12),
                                              cdp=cdp))
                state = cand
                dt = min(2.0 * dt, dt_report - t_local) if cuts else dt
            run.report_states.append(state)
            run.newton.append(its)
        run.final = state
        return run

    def value(self, run: ForwardRun) -> float:
        total = 0.0
        for n, st in enumerate(run.report_states):
            total += float(self.objective(n, st.qws, st.qos, st.bhp))
        return total

    # -- backward --------------------------------------------------------------
    def _states_of_segment(self, run: ForwardRun, k0: int, k1: int) -> List[DeviceOWState]:
        """States at the ends of ministeps k0..k1-1 (recomputed from the
        checkpoint at k0), preceded by the start state of k0."""
        m = self.model
        out = [run.checkpoints[k0]]
        for k in range(k0, k1):
            ms = run.ministeps[k]
            st, ok, _ = self._newton(m, out[-1], ms.dt, ms.cdp)
            if not ok:
                raise RuntimeError(f"recomputing ministep {k} did not converge")
            out.append(st)
        return out

    def gradient(self, run: ForwardRun):
        """(dJ/dln k per cell, dJ/du per report step and well)."""
        m = self.model
        nc, nw = m.nc, m.nw
        n = 2 * nc + 3 * nw
        grad = torch.zeros(nc, dtype=torch.float64, device=self.device)
        gu = np.zeros((len(self.dts), nw))
        lam_next = None
        ms_next = None
        st_next_prev = None
        opts = self.solver.options()
        info = _lib.SolveInfo()
        tmp = torch.empty(n, dtype=torch.float64, device=self.device)
        P = _lib.ptr
        K = len(run.ministeps)
        starts = sorted(run.checkpoints)
        self.adjoint_its = []
        for si in range(len(starts) - 1, -1, -1):
            k0 = starts[si]
            k1 = starts[si + 1] if si + 1 < len(starts) else K
            states = self._states_of_segment(run, k0, k1)
            for k in range(k1 - 1, k0 - 1, -1):
                ms = run.ministeps[k]
                st_prev, st = states[k - k0], states[k - k0 + 1]
                # rhs = -dJ/dx_k - (dR_{k+1}/dx_k)^T lam_{k+1}
                rhs = torch.zeros(n, dtype=torch.float64, device=self.device)
                if ms.ends_report:
                    with torch.enable_grad():   # also inside an autograd backward
                        q = [t.detach().clone().requires_grad_(True) for t in (st.qws, st.qos, st.bhp)]
                        jn = self.objective(ms.report, *q)
                        dq = torch.autograd.grad(jn, q, allow_unused=True)
                    well = rhs[2 * nc:].view(nw, 3)
                    for col, d in enumerate(dq):
                        if d is not None:
                            well[:, col] -= d
                if lam_next is not None:
                    _lib.check(m.lib.pmx_ow_prev_state_transpose(m.h, P(st.p), P(st.sw), float(ms_next.dt),
                                                                  P(lam_next), P(tmp)),
                               "pmx_ow_prev_state_transpose")
                    rhs -= tmp
                m.assemble_device(st_prev, st, ms.dt, ms.cdp)
                ok, pmean = m.converged(ms.dt, st)
                lam = torch.empty(n, dtype=torch.float64, device=self.device)
                rc = m.lib.pmx_ow_adjoint_solve(m.h, ctypes.byref(opts), float(pmean), P(rhs), P(lam),
                                                ctypes.byref(info))
                _lib.check(rc, "pmx_ow_adjoint_solve")
                if rc != 0:
                    raise RuntimeError(f"adjoint solve of ministep {k} did not converge "
                                       f"({info.iterations} its, rel {info.rel_residual:.1e})")
                self.adjoint_its.append(int(info.iterations))
                _lib.check(m.lib.pmx_ow_dlogk(m.h, P(st.p), P(st.sw), P(st.qws), P(st.qos), P(st.bhp),
                                              P(ms.cdp), P(lam), P(grad)), "pmx_ow_dlogk")
                # dR/du = -1 in the control rows (adjoint.py:107-121)
                gu[ms.report] -= lam[2 * nc + 2::3].cpu().numpy()
                lam_next, ms_next = lam, ms
        return grad, gu

    def value_and_gradient(self, logk: np.ndarray):
        run = self.forward(logk)
        J = self.value(run)
        g, gu = self.gradient(run)
        return J, g, gu, run


class LogPermObjective(torch.autograd.Function):
    """J(ln k) as a differentiable torch scalar: the forward runs the
    schedule (keeping checkpoints), the backward runs the adjoint sweep."""

    @staticmethod
    def forward(ctx, logk: torch.Tensor, problem: AdjointProblem):
        run = problem.forward(logk.detach().cpu().numpy())
        ctx.problem, ctx.run = problem, run
        return torch.tensor(problem.value(run), dtype=torch.float64, device=logk.device)

    @staticmethod
    def backward(ctx, gout):
        g, _ = ctx.problem.gradient(ctx.run)
        return gout * g.to(gout.device), None


def log_perm_objective(logk: torch.Tensor, problem: AdjointProblem) -> torch.Tensor:
    return LogPermObjective.apply(logk, problem)
