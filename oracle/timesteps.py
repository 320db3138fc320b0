"""Time-step selectors and the selector-driven report step, restated on the
host for the oracle models (test infrastructure only: tests import this as
the checker; the product never does).

pymrst/modules/ad_core/timesteps/timestep_selectors.py:
  SimpleTimeStepSelectorAD :100-214, FactorTimeStepSelectorAD :217-230,
  IterationCountTimeStepSelectorAD :233-281, StateChangeTimeStepSelectorAD
  :284-394 (the state's own field by name, _get_prop :59-81);
pymrst/modules/ad_core/timesteps/rampup_timesteps.py:19-34 with
pymrst/utils/matlab_sum.py:23-40;
pymrst/modules/ad_core/simulators/nonlinear_solver.py: _solve_adaptive
:770-844 over the oracle model's Newton loop (`model.step`).

Pinned by the reference's own known answers (its __main__ tests of
timestep_selectors.py, simple_time_step_selector.py,
iteration_count_time_step_selector.py, rampup_timesteps.py and
matlab_sum.py) in tests/test_timesteps.py.
"""
import math
from types import SimpleNamespace

import numpy as np


class Simple:
    def __init__(self, max_timestep=math.inf, min_timestep=0.0, max_history=50, max_relative_adjustment=2.0,
                 min_relative_adjustment=0.5, first_rampup_step=math.inf, first_rampup_step_relative=1.0,
                 reset_on_controls_changed=False):
        assert 0.0 < first_rampup_step_relative <= 1.0
        self.max_timestep, self.min_timestep, self.max_history = max_timestep, min_timestep, max_history
        self.max_rel, self.min_rel = max_relative_adjustment, min_relative_adjustment
        self.ramp_abs, self.ramp_rel = first_rampup_step, first_rampup_step_relative
        self.reset_on_change = reset_on_controls_changed
        self.reset()

    def reset(self):
        self.history = []
        self.prev_control = None
        self.start_ctrl, self.first, self.changed, self.limited = True, True, True, True

    def store_timestep(self, rep):
        self.history = (self.history + [rep])[-self.max_history:]

    def new_control_step(self, control=None):
        self.start_ctrl = True
        p = self.prev_control
        pid = getattr(p, "controlId", None) if not isinstance(p, dict) else p.get("controlId")
        cid = getattr(control, "controlId", None) if not isinstance(control, dict) else control.get("controlId")
        differs = (pid != cid) if (pid is not None and cid is not None) else (p is not control)
        if p is None or differs:
            if self.reset_on_change:
                self.reset()
            self.changed = True
            self.prev_control = control
        else:
            self.changed = False

    def pick_timestep(self, dt_prev, dt, model, solver, state_prev, state_curr):
        if self.changed and (self.reset_on_change or self.first):
            dt = min(min(dt, self.ramp_rel * dt), self.ramp_abs)
            self.limited = True
        dt_new = self.compute_timestep(dt, dt_prev, model, solver, state_prev, state_curr)
        f = dt_new / dt
        if not self.start_ctrl:
            f = max(min(f, self.max_rel), self.min_rel)
        out = max(self.min_timestep, min(self.max_timestep, dt * f))
        self.limited = out != dt_new
        self.start_ctrl = self.first = self.changed = False
        return out

    def cut_timestep(self, dt_prev, dt, model, solver, state_prev, state_curr):
        return dt * self.min_rel

    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        return dt


class Factor(Simple):
    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        return dt * self.max_rel


class IterationCount(Simple):
    def __init__(self, target_iteration_count=5, iteration_offset=5, **kw):
        super().__init__(**kw)
        self.target, self.offset = target_iteration_count, iteration_offset

    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        h = self.history
        if not h or not h[-1].converged or (self.changed and self.reset_on_change):
            return dt
        m = solver.max_iterations + self.offset
        tol = (self.target + self.offset) / m
        le1 = (h[-1].iterations + self.offset) / m
        if len(h) < 2 or self.limited or not h[-2].converged:
            return (tol / le1) * h[-1].dt
        le0 = (h[-2].iterations + self.offset) / m
        return (h[-1].dt / h[-2].dt) * (tol * le0 / le1 ** 2) * h[-1].dt


class StateChange(IterationCount):
    def __init__(self, target_props=(), target_change_rel=None, target_change_abs=None, relax_factor=1.0, **kw):
        super().__init__(**kw)
        n = len(target_props)
        self.props = list(target_props)
        self.t_rel = list(target_change_rel) if target_change_rel is not None else [math.inf] * n
        self.t_abs = list(target_change_abs) if target_change_abs is not None else [math.inf] * n
        self.w = relax_factor
        self.prev_state, self.prev_dt = None, math.inf

    def reset(self):
        super().reset()
        self.prev_state, self.prev_dt = None, math.inf

    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        dt0 = dt
        dt = super().compute_timestep(dt, dt_prev, model, solver, state_prev, state_curr)
        if state_prev is None:
            if self.prev_state is None:
                self.prev_state, self.prev_dt = state_curr, dt0
                return dt
            state_prev, dt_prev = self.prev_state, self.prev_dt
        self.prev_state, self.prev_dt = state_curr, dt0
        dtp = math.nan if dt_prev is None else float(dt_prev)
        nxt = math.inf
        for name, tr, ta in zip(self.props, self.t_rel, self.t_abs):
            cur = np.atleast_1d(np.asarray(getattr(state_curr, name), float))
            prv = np.atleast_1d(np.asarray(getattr(state_prev, name), float))
            d = np.abs(cur - prv)
            with np.errstate(divide="ignore", invalid="ignore"):
                rel = d / np.abs(prv)
                rel[np.isnan(rel)] = 0.0
                fa = np.float64((1.0 + self.w) * ta) / np.float64((d.max() if d.size else 0.0) + self.w * ta)
                fr = np.float64((1.0 + self.w) * tr) / np.float64((rel.max() if rel.size else 0.0) + self.w * tr)
            fr = math.inf if math.isnan(fr) else float(fr)
            nxt = float(np.nanmin([nxt, dtp * fr, dtp * float(fa)]))
            dt = min(dt, nxt)
        return dt


def matlab_sum(x):
    v = [float(a) for a in np.asarray(x, dtype=float).ravel()]
    if len(v) <= 43:
        t = 0.0
        for a in v:
            t += a
        return t
    lanes = [0.0] * 4
    n = len(v) // 4 * 4
    for i in range(n):
        lanes[i % 4] += v[i]
    for i in range(n, len(v)):
        lanes[i - n] += v[i]
    return ((lanes[0] + lanes[1]) + lanes[2]) + lanes[3]


def rampup_timesteps(time, dt, n=8):
    if time == 0:
        return np.zeros(0)
    first = dt / 2.0 ** np.array([n] + list(range(n, 0, -1)), dtype=float)
    cs = np.cumsum(first)
    if np.any(cs > time):
        first = first[cs < time]
    rem = np.full(int(np.floor((time - matlab_sum(first)) / dt)), float(dt))
    last = time - matlab_sum(first) - matlab_sum(rem)
    return np.concatenate([first, rem, np.array([last]) if last > 0 else np.zeros(0)])


def solve_adaptive(model, state0, dt, selector, max_cuts=6, control=None):
    """_solve_adaptive over model.step.  Returns (state, converged,
    iterations, cuts, ministeps)."""
    solver = SimpleNamespace(max_iterations=model.max_iterations)
    selector.new_control_step(control)
    dt_min = dt / 2.0 ** max_cuts
    st, prev, dt_prev = state0, None, None
    t, dt_try, failed = 0.0, dt, False
    its = cuts = 0
    taken = []
    while t < dt * (1.0 - 1e-12):
        pick = selector.cut_timestep if failed else selector.pick_timestep
        dt_sel = pick(dt_prev, dt_try, model, solver, prev, st)
        remaining = dt - t
        final = dt_sel >= remaining * (1.0 - 1e-12)
        dt_try = remaining if final else dt_sel
        new, ok, k = model.step(st, dt_try)
        its += k
        if not final or dt_try > 0.9 * dt_sel:
            selector.store_timestep(SimpleNamespace(converged=bool(ok), iterations=int(k), dt=dt_try))
        if ok:
            prev, st, dt_prev = st, new, dt_try
            t += dt_try
            taken.append(dt_try)
            failed = False
        else:
            cuts += 1
            failed = True
            if dt_try <= dt_min:
                return st, False, its, cuts, taken
    return st, True, its, cuts, taken


def simulate_schedule_adaptive(model, state0, dts, selector, max_cuts=6):
    """Report steps through solve_adaptive, stopping at the first failure.
    Returns (states, iterations per step, ministeps per step)."""
    st, states, its, ministeps = state0, [], [], []
    for dt in np.atleast_1d(dts):
        st, ok, k, _, taken = solve_adaptive(model, st, float(dt), selector, max_cuts)
        states.append(st)
        its.append(k)
        ministeps.append(taken)
        if not ok:
            break
    return states, np.array(its), ministeps
