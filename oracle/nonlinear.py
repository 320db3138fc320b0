"""NonLinearSolver's report-step driver restated over an oracle model's
Newton loop (test infrastructure only).

pymrst/modules/ad_core/simulators/nonlinear_solver.py: solve_timestep
:621-702 (the whole dt first, then 2^k equal sub-steps from state0 for
k = 1 .. max_timestep_cuts; iterations of a cut step are the sum over its
sub-steps), _iter_timesteps :955-977 (stop at the first report step that
does not converge).  The ministep is the model's own `step` (the Newton
loop of solve_ministep :239-381 without relaxation or line search).
"""
import numpy as np


def solve_timestep(model, state0, dt, max_cuts=6):
    """Returns (state, converged, iterations, cuts)."""
    st, ok, its = model.step(state0, dt)
    if ok:
        return st, True, its, 0
    for cut in range(1, max_cuts + 1):
        n_sub = 2 ** cut
        st, ok, total = state0, True, 0
        for _ in range(n_sub):
            cand, good, k = model.step(st, dt / n_sub)
            total += k
            if not good:
                ok = False
                break
            st = cand
        if ok:
            return st, True, total, cut
    return st, False, its, max_cuts


def simulate_schedule(model, state0, dts, max_cuts=6):
    """Returns (states, iterations per step, cuts per step)."""
    st = state0
    states, its, cuts = [], [], []
    for dt in np.atleast_1d(dts):
        st, ok, k, c = solve_timestep(model, st, float(dt), max_cuts)
        states.append(st)
        its.append(k)
        cuts.append(c)
        if not ok:
            break
    return states, np.array(its), np.array(cuts)
