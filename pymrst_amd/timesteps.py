"""Time-step selectors for NonLinearSolver (host control logic).

Reference: pymrst/modules/ad_core/timesteps/timestep_selectors.py
(SimpleTimeStepSelectorAD :100-214, FactorTimeStepSelectorAD :217-230,
IterationCountTimeStepSelectorAD :233-281, StateChangeTimeStepSelectorAD
:284-394, _control_changed :84-97, _get_prop :59-81) and
rampup_timesteps.py:19-34 with pymrst/utils/matlab_sum.py.

The selectors see only ministep reports (converged, iterations, dt) and, for
the state-change rule, the largest change of named state fields between two
ministeps.  States are the models' device states: the change is reduced on
the GPU and only the two maxima cross to the host.  Names accepted by the
state-change rule: "pressure", "s" (all phase saturations), "sW", "sG",
"rs" (whichever the state holds).
"""
from __future__ import annotations

import math
from typing import Any, List, Optional, Sequence

import numpy as np

__all__ = ["SimpleTimeStepSelectorAD", "FactorTimeStepSelectorAD", "IterationCountTimeStepSelectorAD",
           "StateChangeTimeStepSelectorAD", "rampup_timesteps", "rampupTimesteps", "matlab_sum"]


def _control_changed(previous: Any, control: Any) -> bool:
    """Controls carrying a `controlId` compare by id, others by identity."""
    def cid(c):
        return c.get("controlId") if isinstance(c, dict) else getattr(c, "controlId", None)
    a, b = cid(previous), cid(control)
    if a is not None and b is not None:
        return a != b
    return previous is not control


class SimpleTimeStepSelectorAD:
    """Hard limits, the first-step ramp-up and the bound on the change of dt
    between ministeps; no rule of its own (`compute_timestep` returns dt)."""

    def __init__(self, max_timestep: float = math.inf, min_timestep: float = 0.0, max_history: int = 50,
                 max_relative_adjustment: float = 2.0, min_relative_adjustment: float = 0.5,
                 first_rampup_step: float = math.inf, first_rampup_step_relative: float = 1.0,
                 reset_on_controls_changed: bool = False, verbose: bool = False):
        if not 0.0 < first_rampup_step_relative <= 1.0:
            raise ValueError("first_rampup_step_relative must be in (0, 1]")
        self.max_timestep = max_timestep
        self.min_timestep = min_timestep
        self.max_history = max_history
        self.max_relative_adjustment = max_relative_adjustment
        self.min_relative_adjustment = min_relative_adjustment
        self.first_rampup_step = first_rampup_step
        self.first_rampup_step_relative = first_rampup_step_relative
        self.reset_on_controls_changed = reset_on_controls_changed
        self.verbose = verbose
        self.reset()

    def reset(self) -> None:
        self.history: List[Any] = []
        self.previous_control: Optional[Any] = None
        self.is_start_of_ctrl_step = True
        self.is_first_step = True
        self.controls_changed = True
        # whether the last dt was set by a limit rather than by the rule
        self.step_limited_by_hard_limits = True

    def store_timestep(self, report: Any) -> None:
        self.history.append(report)
        if len(self.history) > self.max_history:
            del self.history[:len(self.history) - self.max_history]

    def new_control_step(self, control: Any = None) -> None:
        self.is_start_of_ctrl_step = True
        if self.previous_control is None or _control_changed(self.previous_control, control):
            if self.reset_on_controls_changed:
                self.reset()
            self.controls_changed = True
            self.previous_control = control
        else:
            self.controls_changed = False

    def pick_timestep(self, dt_prev: Optional[float], dt: float, model: Any, solver: Any,
                      state_prev: Any, state_curr: Any) -> float:
        if self.controls_changed and (self.reset_on_controls_changed or self.is_first_step):
            dt = min(dt, self.first_rampup_step_relative * dt, self.first_rampup_step)
            self.step_limited_by_hard_limits = True
        dt0 = dt
        dt_new = self.compute_timestep(dt, dt_prev, model, solver, state_prev, state_curr)
        change = dt_new / dt
        if not self.is_start_of_ctrl_step:
            change = max(min(change, self.max_relative_adjustment), self.min_relative_adjustment)
        dt = max(self.min_timestep, min(self.max_timestep, dt * change))
        if self.verbose and dt != dt0:
            print(f"time step changed by a factor {dt / dt0:1.2f}: {dt0:.4g} s -> {dt:.4g} s")
        self.step_limited_by_hard_limits = dt != dt_new
        self.is_start_of_ctrl_step = False
        self.is_first_step = False
        self.controls_changed = False
        return dt

    def cut_timestep(self, dt_prev: Optional[float], dt: float, model: Any, solver: Any,
                     state_prev: Any, state_curr: Any) -> float:
        return dt * self.min_relative_adjustment

    def compute_timestep(self, dt: float, dt_prev: Optional[float], model: Any, solver: Any,
                         state_prev: Any, state_curr: Any) -> float:
        return dt


class FactorTimeStepSelectorAD(SimpleTimeStepSelectorAD):
    """Grow by `max_relative_adjustment` every ministep."""

    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        return dt * self.max_relative_adjustment


class IterationCountTimeStepSelectorAD(SimpleTimeStepSelectorAD):
    """Steer the Newton count toward `target_iteration_count` with the
    response (its + offset) / (max_iterations + offset): one-point update
    after a limited or failed step, two-point update otherwise."""

    def __init__(self, target_iteration_count: int = 5, iteration_offset: int = 5, **kwargs):
        super().__init__(**kwargs)
        self.target_iteration_count = target_iteration_count
        self.iteration_offset = iteration_offset

    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        hist = self.history
        if not hist or not hist[-1].converged or (self.controls_changed and self.reset_on_controls_changed):
            return dt
        off = self.iteration_offset
        denom = solver.max_iterations + off
        target = (self.target_iteration_count + off) / denom
        r1 = (hist[-1].iterations + off) / denom
        dt1 = hist[-1].dt
        if len(hist) < 2 or self.step_limited_by_hard_limits or not hist[-2].converged:
            return (target / r1) * dt1
        r0 = (hist[-2].iterations + off) / denom
        return (dt1 / hist[-2].dt) * (target * r0 / r1 ** 2) * dt1


_FIELDS = {"pressure": ("p", "pressure"), "sW": ("sw", "sW"), "s": ("s",), "rs": ("rs",)}


def _state_field(state: Any, name: str):
    """The named field of a device (DeviceOWState / DeviceBOState) or host
    state; sW and sG of a black-oil state are columns of s."""
    for attr in _FIELDS.get(name, (name,)):
        if hasattr(state, attr):
            return getattr(state, attr)
    if name in ("sW", "sG") and hasattr(state, "s"):
        return state.s[:, 0 if name == "sW" else 2]
    raise KeyError(f"{name!r} is not a field of {type(state).__name__}")


def _max_changes(curr, prev):
    """(max |curr - prev|, max |curr - prev| / |prev|), 0/0 counted as 0 and
    x/0 as infinite; one device-to-host transfer for both."""
    try:
        import torch
        if isinstance(curr, torch.Tensor):
            d = (curr - prev).abs().reshape(-1)
            if d.numel() == 0:
                return 0.0, 0.0
            rel = d / prev.abs().reshape(-1)
            rel = torch.nan_to_num(rel, nan=0.0, posinf=math.inf)
            m = torch.stack([d.max(), rel.max()]).cpu().tolist()
            return float(m[0]), float(m[1])
    except ImportError:
        pass
    d = np.abs(np.asarray(curr, float) - np.asarray(prev, float)).ravel()
    if d.size == 0:
        return 0.0, 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = d / np.abs(np.asarray(prev, float)).ravel()
    rel[np.isnan(rel)] = 0.0
    return float(d.max()), float(rel.max())


class StateChangeTimeStepSelectorAD(IterationCountTimeStepSelectorAD):
    """The iteration-count rule, bounded by target changes of named state
    fields per ministep: dt_prev * (1 + w) T / (change + w T) for each
    absolute and relative target T, the smallest candidate winning."""

    def __init__(self, target_props: Sequence[str] = (), target_change_rel: Optional[Sequence[float]] = None,
                 target_change_abs: Optional[Sequence[float]] = None, relax_factor: float = 1.0, **kwargs):
        super().__init__(**kwargs)
        n = len(target_props)
        self.target_props = list(target_props)
        self.target_change_rel = list(target_change_rel) if target_change_rel is not None else [math.inf] * n
        self.target_change_abs = list(target_change_abs) if target_change_abs is not None else [math.inf] * n
        if len(self.target_change_rel) != n or len(self.target_change_abs) != n:
            raise ValueError("target_change_* must match target_props in length")
        self.relax_factor = relax_factor
        self.previous_state = None
        self.previous_dt = math.inf

    def reset(self) -> None:
        super().reset()
        self.previous_state = None
        self.previous_dt = math.inf

    def compute_timestep(self, dt, dt_prev, model, solver, state_prev, state_curr):
        dt0 = dt
        dt = super().compute_timestep(dt, dt_prev, model, solver, state_prev, state_curr)
        if state_prev is None:
            if self.previous_state is None:
                self.previous_state, self.previous_dt = state_curr, dt0
                return dt
            state_prev, dt_prev = self.previous_state, self.previous_dt
        self.previous_state, self.previous_dt = state_curr, dt0
        w = self.relax_factor
        dt_p = math.nan if dt_prev is None else float(dt_prev)
        dt_next = math.inf
        for name, t_rel, t_abs in zip(self.target_props, self.target_change_rel, self.target_change_abs):
            max_abs, max_rel = _max_changes(_state_field(state_curr, name), _state_field(state_prev, name))
            with np.errstate(divide="ignore", invalid="ignore"):
                f_abs = float(np.float64((1.0 + w) * t_abs) / np.float64(max_abs + w * t_abs))
                f_rel = float(np.float64((1.0 + w) * t_rel) / np.float64(max_rel + w * t_rel))
            if math.isnan(f_rel):
                f_rel = math.inf
            cands = [c for c in (dt_next, dt_p * f_rel, dt_p * f_abs) if not math.isnan(c)]
            dt_next = min(cands) if cands else math.nan
            if self.verbose:
                print(f"  {name!r}: rel {max_rel:.3g} [{t_rel:.3g}], abs {max_abs:.3g} [{t_abs:.3g}]")
            dt = min(dt, dt_next)
        return dt


def matlab_sum(x) -> float:
    """Sum in MATLAB's order of additions (pymrst/utils/matlab_sum.py:23-40):
    one after another up to 43 elements, else four interleaved running sums
    (element i to lane i mod 4) added in order at the end."""
    v = np.asarray(x, dtype=float).ravel().tolist()
    if len(v) <= 43:
        total = 0.0
        for a in v:
            total += a
        return total
    lanes = [0.0, 0.0, 0.0, 0.0]
    full = len(v) - len(v) % 4
    for i in range(full):
        lanes[i % 4] += v[i]
    for i, a in enumerate(v[full:]):
        lanes[i] += a
    total = 0.0
    for a in lanes:
        total += a
    return total


def rampup_timesteps(time: float, dt: float, n: int = 8) -> np.ndarray:
    """MRST's rampupTimesteps: dt / 2^n twice, then doubling up to dt, full
    steps of dt, and the remainder (rampup_timesteps.py:19-34)."""
    if time == 0:
        return np.zeros(0)
    ramp = dt / 2.0 ** np.array([n] + list(range(n, 0, -1)), dtype=float)
    ramp = ramp[np.cumsum(ramp) < time] if np.any(np.cumsum(ramp) > time) else ramp
    left = time - matlab_sum(ramp)
    full = np.full(int(np.floor(left / dt)), float(dt))
    last = time - matlab_sum(ramp) - matlab_sum(full)
    return np.concatenate([ramp, full, [last] if last > 0 else []])


rampupTimesteps = rampup_timesteps
