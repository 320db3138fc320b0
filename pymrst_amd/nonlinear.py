"""NonLinearSolver for the device models: the reference's Newton driver with
time-step cutting, over the pmx models' device Newton primitives.

Reference: pymrst/modules/ad_core/simulators/nonlinear_solver.py
(StepReport :52-89, SimulationReport :92-110, NonLinearSolver.__init__
:117-160, solve_ministep :239-381, solve_timestep :621-702,
_solve_adaptive :770-844, simulate_schedule :847-890, _iter_timesteps
:955-977).  Time-step selectors: `pymrst_amd.timesteps`.

The models (`pymrst_amd.ow.TwoPhaseOilWaterModelAD`,
`pymrst_amd.bo.ThreePhaseBlackOilModelAD`) provide three device steps in
place of the reference's get_equations / get_convergence / _solve_linear +
update_state:

    model.begin_ministep(state0)              per ministep, before Newton
    model.linearize(state0, state, dt)       -> (converged, residuals dict)
    model.newton_update(state)               -> False when the linear solve failed

States stay on the device between calls.  Relaxation, line search and the
time-step selectors of the reference are not part of this driver; a
structured Schedule (control periods) is not either: `schedule` is a
sequence of report-step lengths.
"""
from dataclasses import dataclass, field
from time import perf_counter
from typing import Dict, List, Sequence

import numpy as np


@dataclass
class StepReport:
    converged: bool
    iterations: int
    dt: float
    residuals: List[Dict[str, float]] = field(default_factory=list)
    cuts: int = 0
    ministeps: List[float] = field(default_factory=list)       # the ministeps a selector took, summing to dt
    linear_solves: List[tuple] = field(default_factory=list)   # (iterations, relative residual) per solve
    wall_seconds: float = 0.0
    assembly_seconds: float = 0.0


@dataclass
class SimulationReport:
    steps: List[StepReport] = field(default_factory=list)

    @property
    def converged(self) -> bool:
        return all(s.converged for s in self.steps)

    @property
    def total_iterations(self) -> int:
        return sum(s.iterations for s in self.steps)

    @property
    def total_cuts(self) -> int:
        return sum(s.cuts for s in self.steps)


class NonLinearSolver:
    """Newton with time-step cutting (the reference's defaults: 25
    iterations, 6 cuts).  With a `timestep_selector` (one of
    `pymrst_amd.timesteps`) a report step is crossed in the ministeps the
    selector picks, the last one trimmed to land on the report time.
    `observer(event, payload)` receives the reference's event names:
    report_step_started, ministep_started, nonlinear_iteration_completed,
    linear_solve_failed, timestep_cut, step_accepted, step_rejected."""

    def __init__(self, max_iterations: int = 25, max_timestep_cuts: int = 6, verbose: bool = False,
                 observer=None, timestep_selector=None):
        self.max_iterations = max_iterations
        self.max_timestep_cuts = max_timestep_cuts
        self.timestep_selector = timestep_selector
        self.verbose = verbose
        self.observer = observer
        self.observer_failures: List[dict] = []
        self._context: dict = {}

    def _observe(self, event: str, **payload):
        if self.observer is None:
            return
        try:
            self.observer(event, {**self._context, **payload})
        except Exception as error:   # an observer never stops the run
            self.observer_failures.append({"event": event, "type": type(error).__name__, "message": str(error)})

    # -- one Newton loop at a fixed dt (solve_ministep :239-381) ---------------
    def solve_ministep(self, model, state0, state, dt: float):
        report = StepReport(converged=False, iterations=0, dt=dt)
        self._observe("ministep_started", dt=float(dt))
        started = perf_counter()
        st = state
        model.begin_ministep(state0)
        for it in range(self.max_iterations):
            t0 = perf_counter()
            converged, res = model.linearize(state0, st, dt)
            report.assembly_seconds += perf_counter() - t0
            report.residuals.append(res)
            report.iterations = it
            self._observe("nonlinear_iteration_completed", iteration=int(it), dt=float(dt),
                          converged=bool(converged), residuals=dict(res))
            if converged:
                report.converged = True
                report.wall_seconds = perf_counter() - started
                return st, report
            ok = model.newton_update(st)
            report.linear_solves.append(model.last_linear_solve)
            if not ok:
                # an iterative solve that fails is a time-step failure: the
                # caller cuts dt and retries from the last converged state
                report.residuals[-1] = dict(report.residuals[-1], linear_failure=float("inf"))
                self._observe("linear_solve_failed", iteration=int(it), dt=float(dt),
                              linear_iterations=int(model.last_linear_solve[0]))
                report.wall_seconds = perf_counter() - started
                return st, report
        t0 = perf_counter()
        converged, res = model.linearize(state0, st, dt)
        report.assembly_seconds += perf_counter() - t0
        report.residuals.append(res)
        self._observe("nonlinear_iteration_completed", iteration=int(self.max_iterations), dt=float(dt),
                      converged=bool(converged), residuals=dict(res))
        report.converged = bool(converged)
        report.iterations = self.max_iterations
        report.wall_seconds = perf_counter() - started
        return st, report

    # -- one report step with cutting (solve_timestep :621-702) ------------------
    def solve_timestep(self, model, state0, dt: float, control=None):
        """Take a step of dt; on failure retry as 2^k equal sub-steps from
        state0, k = 1 .. max_timestep_cuts."""
        if control is not None:
            raise NotImplementedError("control periods are not supported by the device driver")
        state0 = model.device_state(state0)
        if self.timestep_selector is not None:
            return self._solve_adaptive(model, state0, dt, control)
        self._context.update(attempt_index=0, ministep_index=0)
        st, report = self.solve_ministep(model, state0, state0.copy(), dt)
        if report.converged:
            return st, report
        last_failure = report
        for cut in range(1, self.max_timestep_cuts + 1):
            n_sub = 2 ** cut
            dt_sub = dt / n_sub
            self._observe("timestep_cut", cut=int(cut), subdivisions=int(n_sub), report_dt=float(dt),
                          trial_dt=float(dt_sub))
            if self.verbose:
                print(f"  step failed; cutting to {n_sub} x {dt_sub:.3g} s")
            st, ok, iters = state0, True, 0
            residuals, linear = [], []
            for substep in range(n_sub):
                self._context.update(attempt_index=cut, ministep_index=substep)
                st_new, rep = self.solve_ministep(model, st, st.copy(), dt_sub)
                iters += rep.iterations
                residuals.extend(rep.residuals)
                linear.extend(rep.linear_solves)
                if not rep.converged:
                    ok = False
                    last_failure = rep
                    break
                st = st_new
            if ok:
                return st, StepReport(converged=True, iterations=iters, dt=dt, cuts=cut, residuals=residuals,
                                      linear_solves=linear)
        return st, StepReport(converged=False, iterations=report.iterations, dt=dt, cuts=self.max_timestep_cuts,
                              residuals=list(last_failure.residuals),
                              linear_solves=list(last_failure.linear_solves))

    # -- one report step in selector-chosen ministeps (_solve_adaptive :770-844)
    def _solve_adaptive(self, model, state0, dt: float, control=None):
        """Ministeps from the selector (its cut rule after a failure); the
        last one is trimmed to the report time and, when trimmed below 0.9
        of the selector's pick, kept out of its history.  A failure at or
        below dt / 2^max_timestep_cuts ends the report step."""
        sel = self.timestep_selector
        sel.new_control_step(control)
        dt_min = dt / 2.0 ** self.max_timestep_cuts
        st, state_prev, dt_prev = state0, None, None
        t_local, dt_try, failed = 0.0, dt, False
        iters = cuts = attempt = 0
        taken: List[float] = []
        residuals, linear = [], []
        while t_local < dt * (1.0 - 1e-12):
            pick = sel.cut_timestep if failed else sel.pick_timestep
            dt_sel = pick(dt_prev, dt_try, model, self, state_prev, st)
            remaining = dt - t_local
            is_final = dt_sel >= remaining * (1.0 - 1e-12)
            dt_try = remaining if is_final else dt_sel
            self._context.update(attempt_index=attempt, ministep_index=len(taken))
            st_new, rep = self.solve_ministep(model, st, st.copy(), dt_try)
            attempt += 1
            iters += rep.iterations
            residuals.extend(rep.residuals)
            linear.extend(rep.linear_solves)
            if not is_final or dt_try > 0.9 * dt_sel:
                sel.store_timestep(rep)
            if rep.converged:
                state_prev, st = st, st_new
                dt_prev = dt_try
                t_local += dt_try
                taken.append(dt_try)
                failed = False
            else:
                cuts += 1
                failed = True
                if dt_try <= dt_min:
                    if self.verbose:
                        print(f"  ministep of {dt_try:.3g} s failed at the floor ({dt_min:.3g} s)")
                    return st, StepReport(converged=False, iterations=iters, dt=dt, cuts=cuts, ministeps=taken,
                                          residuals=residuals, linear_solves=linear)
        return st, StepReport(converged=True, iterations=iters, dt=dt, cuts=cuts, ministeps=taken,
                              residuals=residuals, linear_solves=linear)

    # -- a schedule of report steps (simulate_schedule :847-890) ----------------
    def simulate_schedule(self, model, state0, schedule: Sequence[float], after_step_fn=None,
                          keep_device: bool = False):
        """Run the report steps in `schedule`; stop at the first step that
        does not converge (as _iter_timesteps :955-977).  `after_step_fn`
        is called as fn(model, states, reports, solver, schedule, simtime)
        and returns (model, states, reports, solver, ok); ok = False stops
        the run.  Returns (states, SimulationReport); states are host states
        unless keep_device."""
        states, report = [], SimulationReport()
        simtime: List[float] = []
        elapsed = 0.0
        st = model.device_state(state0)
        for index, dt in enumerate(np.atleast_1d(np.asarray(schedule, dtype=float))):
            self._context = {"report_step_index": int(index)}
            self._observe("report_step_started", dt=float(dt))
            started = perf_counter()
            st, rep = self.solve_timestep(model, st, float(dt))
            rep.wall_seconds = perf_counter() - started
            states.append(st if keep_device else st.to_host())
            report.steps.append(rep)
            elapsed += float(dt)
            simtime.append(elapsed)
            if self.verbose:
                print(f"Report step {index + 1}: {'accepted' if rep.converged else 'FAILED'}, "
                      f"{rep.iterations} Newton iterations, {rep.cuts} cuts")
            # the reference hands each report step (a failed one too) to the
            # hook before the step's accept/reject event
            if after_step_fn is not None:
                model, states, steps, _, ok = after_step_fn(model, states, report.steps, self, schedule, simtime)
                report.steps = list(steps)
                if not ok:
                    break
            self._observe("step_accepted" if rep.converged else "step_rejected", dt=float(dt),
                          iterations=int(rep.iterations), cuts=int(rep.cuts))
            if not rep.converged:
                break
        return states, report
