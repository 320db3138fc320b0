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


class StabilisedNewton:
    """The reference's NonLinearSolver.solve_ministep with its stabilisation
    (simulators/nonlinear_solver.py:239-381): heuristic relaxation
    (_update_relaxation :627-632), the back-tracking line search
    (_line_search :634-648) taken only while `convergence_issues` (or
    always_use_stabilization) holds, and the per-equation oscillation /
    stagnation diagnosis that sets it (:427-545), over an oracle model
    (residual_and_jacobian, converged, solve_increment, apply_update,
    compute_cdp_per_well).  Residuals are named CNV and MB."""

    def __init__(self, max_iterations=25, use_relaxation=False, use_line_search=False, relaxation_min=0.5,
                 relaxation_increment=0.1, line_search_max=5, always_use_stabilization=False,
                 oscillation_threshold=1.0, stagnate_tol=1e-2):
        self.max_iterations = max_iterations
        self.use_relaxation, self.use_line_search = use_relaxation, use_line_search
        self.relaxation_min, self.relaxation_increment = relaxation_min, relaxation_increment
        self.line_search_max = line_search_max
        self.always_use_stabilization = always_use_stabilization
        self.oscillation_threshold, self.stagnate_tol = oscillation_threshold, stagnate_tol
        self.issues = False

    @staticmethod
    def _residuals(model, eqs, dt, st):
        ok, (cnv, mb) = model.converged(eqs, dt, st)
        return ok, {"CNV": float(cnv), "MB": float(mb)}

    def _diagnose(self, model, history, it):
        names = list(history[-1].keys())
        h = np.array([[r[n] for n in names] for r in history])
        tol = {"CNV": model.tol_cnv, "MB": model.tol_mb}
        done = np.array([history[-1][n] <= tol[n] for n in names])
        if done.all():
            self.issues = False
            return
        osc = np.zeros(len(names), dtype=bool)
        if it >= 2:
            with np.errstate(divide="ignore", invalid="ignore"):
                q = (h[it] - h[it - 1]) / (h[it - 1] - h[it - 2])
            osc = np.where(np.isfinite(q), q < 0.0, False)
        stag = np.zeros(len(names), dtype=bool)
        if it >= 1:
            with np.errstate(divide="ignore", invalid="ignore"):
                ch = np.abs(h[it] - h[it - 1]) / h[it - 1]
            stag = np.where(np.isfinite(ch), ch < self.stagnate_tol, False)
        bad = (osc | stag) & ~done
        self.issues = bool(bad.sum() >= self.oscillation_threshold * (~done).sum())

    def step(self, model, state0, dt):
        """One ministep: (state, converged, iterations, line-search trials)."""
        st = state0.copy()
        cdp = model.compute_cdp_per_well(state0)
        relax, prev = 1.0, np.inf
        self.issues = False
        history, searches = [], 0
        stab = self.use_relaxation or self.use_line_search or self.always_use_stabilization
        for it in range(self.max_iterations):
            r, J, eqs = model.residual_and_jacobian(state0, st, dt, cdp)
            ok, res = self._residuals(model, eqs, dt, st)
            history.append(res)
            norm = max(res.values())
            if ok:
                if hasattr(model, "after_convergence"):     # nonlinear_solver.py:282-288
                    st = model.after_convergence(st, state0, dt)
                return st, True, it, searches
            struggling = self.always_use_stabilization or self.issues
            if self.use_relaxation:
                relax = (max(relax - self.relaxation_increment, self.relaxation_min) if norm > prev
                         else min(relax + self.relaxation_increment, 1.0))
            prev = norm
            upd = model.solve_increment(r, J, st)
            base = st.copy()
            model.apply_update(st, relax * upd)
            if self.use_line_search and struggling:
                alpha = relax
                for i in range(self.line_search_max):
                    _, _, e2 = model.residual_and_jacobian(state0, st, dt, cdp)
                    if max(self._residuals(model, e2, dt, st)[1].values()) <= norm:
                        searches += i
                        break
                    alpha *= 0.5
                    st = base.copy()
                    model.apply_update(st, alpha * upd)
                else:
                    searches += self.line_search_max
            if stab:
                self._diagnose(model, history, it)
        self.issues = False
        r, J, eqs = model.residual_and_jacobian(state0, st, dt, cdp)
        return st, self._residuals(model, eqs, dt, st)[0], self.max_iterations, searches


def simulate_stabilised(model, state0, dts, newton: StabilisedNewton, max_cuts=6):
    """simulate_schedule with the stabilised ministep; also returns the
    line-search trials per report step."""

    class _Wrap:
        def __init__(self):
            self.searches = 0

        def step(self, st0, dt):
            st, ok, its, ls = newton.step(model, st0, dt)
            self.searches += ls
            return st, ok, its

    w = _Wrap()
    states, its, cuts, searches = [], [], [], []
    st = state0
    for dt in np.atleast_1d(dts):
        before = w.searches
        st, ok, k, c = solve_timestep(w, st, float(dt), max_cuts)
        states.append(st)
        its.append(k)
        cuts.append(c)
        searches.append(w.searches - before)
        if not ok:
            break
    return states, np.array(its), np.array(cuts), np.array(searches)
