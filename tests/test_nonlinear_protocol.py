"""The device NonLinearSolver's stabilisation protocol against the
reference's own known answers (tests/ad_core/solvers/
test_nonlinear_solver_protocol.py): oscillation and stagnation diagnosis per
equation, the relaxation types, message prefixes and line-search names; and
the oracle's stabilised Newton pinned by the reference's paired waterflood
(inert at its defaults: the golden Newton counts).  CPU only."""
import numpy as np
import pytest

from pymrst_amd.nonlinear import NonLinearSolver


def _solver(**kw):
    s = NonLinearSolver()
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_relaxation_is_inert_at_its_defaults():
    s = _solver()
    assert s.relaxation_parameter == 1.0 and s.relaxation_type == "dampen"
    out, rep = s.stabilize_newton_increments(None, None, [np.array([1.0, 2.0]), np.array([3.0])])
    assert out[0].tolist() == [1.0, 2.0] and out[1].tolist() == [3.0] and rep["relaxation_parameter"] == 1.0


def test_oscillation_diagnosis():
    s = _solver()
    assert s.check_for_oscillations(np.array([[1.0], [0.1], [0.5]]), 2)[0]          # down then up
    assert not s.check_for_oscillations(np.array([[1.0], [0.1], [0.01]]), 2)[0]     # descending
    assert not s.check_for_oscillations(np.array([[0.01], [0.1], [1.0]]), 2)[0]     # rising
    assert not s.check_for_oscillations(np.array([[1.0], [0.5]]), 1).any()          # needs three
    flags = s.check_for_oscillations(np.array([[1.0, 1.0], [0.1, 0.5], [0.5, 0.2]]), 2)
    assert flags[0] and not flags[1]                                                # per equation


def test_stagnation_diagnosis():
    s = _solver(stagnate_tol=1e-2)
    assert s.check_for_stagnation(np.array([[1.0], [1.0 - 1e-6]]), 1)[0]
    assert not s.check_for_stagnation(np.array([[1.0], [0.1]]), 1)[0]
    assert not s.check_for_stagnation(np.array([[1e-8], [0.5e-8]]), 1)[0]           # scale-free
    assert not s.check_for_stagnation(np.array([[1e8], [0.5e8]]), 1)[0]
    assert not _solver().check_for_stagnation(np.array([[1.0]]), 0).any()


def test_relaxation_types():
    s = _solver(relaxation_type="dampen", relaxation_parameter=0.5)
    for _ in range(2):     # twice: dampening never mixes with the previous step
        out, _ = s.stabilize_newton_increments(None, None, [np.array([2.0, 4.0])])
        assert out[0].tolist() == [1.0, 2.0]
    s = _solver(relaxation_type="sor", relaxation_parameter=0.5)
    out, _ = s.stabilize_newton_increments(None, None, [np.array([4.0])])
    assert out[0].tolist() == [4.0]                  # nothing to mix with yet
    out, _ = s.stabilize_newton_increments(None, None, [np.array([0.0])])
    assert out[0].tolist() == [2.0]                  # 0.5 * 0 + 0.5 * 4
    s = _solver(relaxation_type="none", relaxation_parameter=0.1)
    for v in (7.0, 1.0):
        assert s.stabilize_newton_increments(None, None, [np.array([v])])[0][0].tolist() == [v]
    with pytest.raises(ValueError, match="magic"):
        _solver(relaxation_type="magic", relaxation_parameter=0.5).stabilize_newton_increments(
            None, None, [np.array([1.0])])
    s = _solver(relaxation_type="dampen", relaxation_parameter=1.0)
    s.stabilize_newton_increments(None, None, [np.array([3.0])])
    assert s.previous_increment[0].tolist() == [3.0]


def test_names_and_prefixes():
    assert _solver().get_id() == "" and _solver(identifier="outer").get_id() == "outer: "
    s = _solver()
    assert s.get_active_names(["water", "oil", "well"]).tolist() == [True] * 3
    s.linesearch_convergence_names = ["oil"]
    assert s.get_active_names(["water", "oil", "well"]).tolist() == [False, True, False]


def test_relaxation_schedule_and_convergence_issues():
    """Relaxation falls by relaxation_increment while the residual rises,
    not below relaxation_min, and recovers to 1; convergence_issues is set
    when every unconverged equation oscillates or stagnates."""
    s = _solver()
    r = 1.0
    for norm, prev, want in ((2.0, 1.0, 0.9), (3.0, 2.0, 0.8), (4.0, 3.0, 0.7), (5.0, 4.0, 0.6),
                             (6.0, 5.0, 0.5), (7.0, 6.0, 0.5), (1.0, 7.0, 0.6)):
        r = s._update_relaxation(r, norm, prev)
        assert r == pytest.approx(want)
    model = type("M", (), {"tol_cnv": 1e-3, "tol_mb": 1e-7})()
    hist = [{"CNV": 1.0, "MB": 1e-3}, {"CNV": 0.1, "MB": 1e-3 * (1 - 1e-4)}, {"CNV": 0.5, "MB": 1e-3 * (1 - 2e-4)}]
    s._update_convergence_issues(hist, 2, model)
    assert s.convergence_issues                      # CNV oscillates, MB stagnates
    hist[-1] = {"CNV": 0.01, "MB": 1e-4}
    s._update_convergence_issues(hist, 2, model)
    assert not s.convergence_issues


def test_oracle_stabilised_newton_is_inert_at_its_defaults():
    from cases import build_oracle, golden, paired_inputs
    from oracle.nonlinear import StabilisedNewton, simulate_stabilised
    inp = paired_inputs()
    m, s0 = build_oracle(inp)
    states, its, _, searches = simulate_stabilised(m, s0, inp["dts"][:4], StabilisedNewton())
    ref = golden("paired_waterflood.npz")
    assert its.tolist() == ref["newton_iterations"][:4].tolist() and searches.sum() == 0
    assert np.abs(states[-1].pressure - ref["pressure"][3]).max() <= 1e-12 * np.abs(ref["pressure"]).max()
