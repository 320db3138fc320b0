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

    model.apply_increment(state, scale)      the last increment again, scaled
                                             (line search trials)

States stay on the device between calls.  The reference's stabilisation is
here: heuristic relaxation (use_relaxation, _update_relaxation :627-632),
the back-tracking line search (use_line_search, gated like the reference
on `convergence_issues` or always_use_stabilization, :596-614), MRST's
per-equation oscillation and stagnation diagnosis (:427-545) and its
relaxation types (stabilize_newton_increments :547-575).  A structured
Schedule (control periods) is not part of this driver: `schedule` is a
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
    line_searches: int = 0
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

    # residual names -> the model tolerance each is judged by (MRST's per-name
    # ResidualsConverged); unknown names count as not converged
    _TOLERANCE_OF = {"CNV": "tol_cnv", "MB": "tol_mb", "well": "tol_well", "control": "tol_well",
                     "extra": "tol_well"}

    def __init__(self, max_iterations: int = 25, max_timestep_cuts: int = 6, verbose: bool = False,
                 observer=None, timestep_selector=None, use_line_search: bool = False,
                 use_relaxation: bool = False, relaxation_min: float = 0.5, relaxation_increment: float = 0.1,
                 line_search_max: int = 5, always_use_stabilization: bool = False,
                 oscillation_threshold: float = 1.0):
        self.max_iterations = max_iterations
        self.max_timestep_cuts = max_timestep_cuts
        self.timestep_selector = timestep_selector
        self.verbose = verbose
        self.observer = observer
        self.observer_failures: List[dict] = []
        self._context: dict = {}
        self.use_line_search = bool(use_line_search)
        self.use_relaxation = bool(use_relaxation)
        self.relaxation_min = float(relaxation_min)
        self.relaxation_increment = float(relaxation_increment)
        self.line_search_max = int(line_search_max)
        self.always_use_stabilization = bool(always_use_stabilization)
        self.oscillation_threshold = float(oscillation_threshold)
        self.convergence_issues = False
        # MRST's relaxation model: a fixed parameter under a named type, inert at 1
        self.relaxation_type = "dampen"
        self.relaxation_parameter = 1.0
        self.previous_increment = None
        self.stagnate_tol = 1.0e-2
        self.identifier = ""
        self.linesearch_convergence_names = None

    # -- convergence diagnosis (MRST NonLinearSolver.m) -----------------------
    def check_for_oscillations(self, res, index: int) -> np.ndarray:
        """Per equation (column of res, one row per iteration): did the
        residual change direction between iterations index-2, index-1 and
        index?  Needs three iterations."""
        hist = np.atleast_2d(np.asarray(res, dtype=float))
        if index < 2:
            return np.zeros(hist.shape[1], dtype=bool)
        before = hist[index - 1] - hist[index - 2]
        now = hist[index] - hist[index - 1]
        with np.errstate(divide="ignore", invalid="ignore"):
            q = now / before
        return np.isfinite(q) & (q < 0.0)

    def check_for_stagnation(self, res, index: int) -> np.ndarray:
        """Per equation: did the residual move by less than stagnate_tol of
        its previous value?  Needs two iterations."""
        hist = np.atleast_2d(np.asarray(res, dtype=float))
        if index < 1:
            return np.zeros(hist.shape[1], dtype=bool)
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.abs(hist[index] - hist[index - 1]) / hist[index - 1]
        return np.isfinite(rel) & (rel < self.stagnate_tol)

    checkForOscillations = check_for_oscillations
    checkForStagnation = check_for_stagnation

    def _converged_by_name(self, res: Dict[str, float], names, model) -> np.ndarray:
        ok = np.zeros(len(names), dtype=bool)
        for i, name in enumerate(names):
            tol = getattr(model, self._TOLERANCE_OF.get(name, ""), None) if name in self._TOLERANCE_OF else None
            if tol is not None and res.get(name) is not None:
                ok[i] = res[name] <= tol
        return ok

    def _update_convergence_issues(self, residuals: List[Dict[str, float]], index: int, model) -> None:
        """convergence_issues: enough of the unconverged equations oscillate
        or stagnate (at least oscillation_threshold of them); read as the
        next iteration's `struggling` signal (NonLinearSolver.m:402-438)."""
        names = list(residuals[-1].keys())
        hist = np.array([[r.get(n, np.nan) for n in names] for r in residuals], dtype=float)
        ok = self._converged_by_name(residuals[-1], names, model)
        if ok.all():
            self.convergence_issues = False
            return
        bad = self.check_for_oscillations(hist, index) | self.check_for_stagnation(hist, index)
        open_ = ~ok
        self.convergence_issues = bool((bad & open_).sum() >= self.oscillation_threshold * open_.sum())

    def stabilize_newton_increments(self, model, problem, dx):
        """Relax the increment blocks dx by relaxation_type: 'dampen' scales
        them by relaxation_parameter, 'sor' blends them with the previous
        increment, 'none' keeps them; inert at a parameter of 1.  The
        increment is remembered whatever the type.  Returns (dx, report)."""
        w = float(self.relaxation_parameter)
        kind = str(self.relaxation_type).lower()
        if kind not in ("dampen", "sor", "none"):
            raise ValueError(f"relaxation_type {self.relaxation_type!r} is not one of 'dampen', 'sor', 'none'")
        if w < 1.0 and kind == "dampen":
            dx = [w * d for d in dx]
        elif w < 1.0 and kind == "sor" and self.previous_increment is not None:
            dx = [w * d + (1.0 - w) * q for d, q in zip(dx, self.previous_increment)]
        self.previous_increment = list(dx)
        return dx, {"relaxationParameter": w, "relaxation_parameter": w}

    stabilizeNewtonIncrements = stabilize_newton_increments

    def get_id(self) -> str:
        return f"{self.identifier}: " if self.identifier else ""

    def get_active_names(self, names) -> np.ndarray:
        """Which residuals the line search judges by: all, or those named
        in linesearch_convergence_names (case-insensitive)."""
        names = list(names)
        if not self.linesearch_convergence_names:
            return np.ones(len(names), dtype=bool)
        keep = {str(n).lower() for n in self.linesearch_convergence_names}
        return np.array([str(n).lower() in keep for n in names], dtype=bool)

    getId = get_id
    getActiveNames = get_active_names

    def _update_relaxation(self, relax: float, norm: float, prev_norm: float) -> float:
        """Damp by relaxation_increment (to relaxation_min) when the residual
        rose, recover towards 1 when it did not."""
        if norm > prev_norm:
            return max(relax - self.relaxation_increment, self.relaxation_min)
        return min(relax + self.relaxation_increment, 1.0)

    def _line_search(self, model, state0, base, relax: float, dt: float, norm0: float):
        """Halve the step along the last increment from `base` until the
        residual norm is no larger than norm0 (at most line_search_max
        trials).  Returns (state, trials)."""
        alpha = relax
        trial = base.copy()
        model.apply_increment(trial, alpha)
        for i in range(self.line_search_max):
            _, res = model.linearize(state0, trial, dt)
            if max(res.values()) <= norm0:
                return trial, i
            alpha *= 0.5
            trial = base.copy()
            model.apply_increment(trial, alpha)
        return trial, self.line_search_max

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
        relax, prev_norm = 1.0, np.inf
        self.convergence_issues = False
        stabilising = self.use_relaxation or self.use_line_search or self.always_use_stabilization
        model.begin_ministep(state0)
        for it in range(self.max_iterations):
            t0 = perf_counter()
            converged, res = model.linearize(state0, st, dt)
            report.assembly_seconds += perf_counter() - t0
            report.residuals.append(res)
            report.iterations = it
            norm = max(res.values()) if res else 0.0
            self._observe("nonlinear_iteration_completed", iteration=int(it), dt=float(dt),
                          converged=bool(converged), residuals=dict(res), relaxation=float(relax))
            if converged:
                report.converged = True
                if hasattr(model, "after_convergence"):   # nonlinear_solver.py:282-288
                    st = model.after_convergence(st, state0, dt)
                report.wall_seconds = perf_counter() - started
                self.convergence_issues = False
                return st, report
            # the previous iteration's diagnosis decides this one's line search
            struggling = self.always_use_stabilization or self.convergence_issues
            if self.use_relaxation:
                relax = self._update_relaxation(relax, norm, prev_norm)
            prev_norm = norm
            searching = self.use_line_search and struggling
            base = st.copy() if searching else None
            ok = model.newton_update(st, relax) if relax != 1.0 else model.newton_update(st)
            report.linear_solves.append(model.last_linear_solve)
            if not ok:
                # an iterative solve that fails is a time-step failure: the
                # caller cuts dt and retries from the last converged state
                report.residuals[-1] = dict(report.residuals[-1], linear_failure=float("inf"))
                self._observe("linear_solve_failed", iteration=int(it), dt=float(dt),
                              linear_iterations=int(model.last_linear_solve[0]))
                report.wall_seconds = perf_counter() - started
                return st, report
            if searching:
                st, trials = self._line_search(model, state0, base, relax, dt, norm)
                report.line_searches += trials
            if stabilising:
                self._update_convergence_issues(report.residuals, it, model)
        self.convergence_issues = False
        t0 = perf_counter()
        converged, res = model.linearize(state0, st, dt)
        report.assembly_seconds += perf_counter() - t0
        report.residuals.append(res)
        self._observe("nonlinear_iteration_completed", iteration=int(self.max_iterations), dt=float(dt),
                      converged=bool(converged), residuals=dict(res), relaxation=float(relax))
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
