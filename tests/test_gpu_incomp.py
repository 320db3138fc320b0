"""GPU incompTPFA and computeTrans against MRST's fixtures and the oracle
(rtol 1e-9, examples/validation/validate_phase2.py:88-117)."""
import numpy as np
import pytest

import oracle.grid as og
from oracle.incomp import add_bc as o_add_bc
from oracle.incomp import add_source as o_add_source
from oracle.incomp import incomp_tpfa as o_incomp

from cases import golden

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MU1 = 1 * og.CENTI * og.POISE


def close(a, b, rtol, atol_floor=1e-15):
    a = np.ravel(np.asarray(a, float))
    b = np.ravel(np.asarray(b, float))
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    return err <= rtol * scale + atol_floor, err / scale


def _device_case(dims, phys, perm, poro):
    import pymrst_amd as pa
    from pymrst_amd import incomp as pi
    g = pa.cart_grid(dims, phys)
    rock = pa.make_rock(g, perm, poro)
    solver = pi.IncompTPFA(g)
    return pa, pi, g, rock, solver


def test_qfs_matches_mrst():
    pa, pi, g, rock, solver = _device_case([32, 32, 1], [500.0, 500.0, 10.0],
                                           (10 + (7 * np.arange(1024)) % 100) * og.MILLI * og.DARCY, 0.2)
    T = solver.compute_trans(rock)
    src = pi.add_source(pi.add_source(None, [0], 1e-3), [g.num_cells - 1], -1e-3)
    st = solver.solve(pi.init_res_sol(g, 0.0), T, pi.init_single_fluid(MU1, 1000.0), src=src)
    ref = golden("p2_qfs.npz")
    for key, val in (("hT", T.cpu().numpy()), ("p", st.pressure), ("flux", st.flux), ("fp", st.facePressure)):
        ok, e = close(val, ref[key], 1e-9)
        assert ok, f"{key}: rel {e:.2e}"


def test_bc_matches_mrst():
    pa, pi, g, rock, solver = _device_case([20, 10, 5], [200.0, 100.0, 50.0],
                                           (1 + (3 * np.arange(1000)) % 50) * og.MILLI * og.DARCY, 0.25)
    T = solver.compute_trans(rock)
    bd = np.any(g.neighbors == -1, axis=1)
    fc = g.face_centroids
    bc = pi.add_bc(None, np.flatnonzero(bd & (fc[:, 0] < 1e-10)), "pressure", 200 * og.BARSA)
    bc = pi.add_bc(bc, np.flatnonzero(bd & (fc[:, 0] > 200 - 1e-10)), "pressure", 100 * og.BARSA)
    bc = pi.add_bc(bc, np.flatnonzero(bd & (fc[:, 1] < 1e-10)), "flux", 1e-6)
    st = solver.solve(pi.init_res_sol(g, 0.0), T, pi.init_single_fluid(MU1, 1000.0), bc=bc)
    ref = golden("p2_bc.npz")
    for key, val in (("hT", T.cpu().numpy()), ("p", st.pressure), ("flux", st.flux), ("fp", st.facePressure)):
        ok, e = close(val, ref[key], 1e-9)
        assert ok, f"{key}: rel {e:.2e}"


def test_wells_gravity_matches_mrst():
    perm = np.tile(np.array([100.0, 100.0, 10.0]) * og.MILLI * og.DARCY, (300, 1))
    pa, pi, g, rock, solver = _device_case([10, 10, 3], [1000.0, 1000.0, 30.0], perm, 0.3)
    T = solver.compute_trans(rock)
    gv = pi.gravity_vector(3)
    w = pa.add_well(None, g, rock, [0, 100, 200], "rate", 5e-3, compi=[1.0], gravity=gv)
    w = pa.add_well(w, g, rock, [99, 199, 299], "bhp", 100 * og.BARSA, compi=[1.0], gravity=gv)
    st = solver.solve(pi.init_res_sol(g, 0.0), T, pi.init_single_fluid(MU1, 1000.0), wells=w, gravity=gv)
    ref = golden("p2_wells.npz")
    pairs = (("hT", T.cpu().numpy()), ("p", st.pressure), ("flux", st.flux),
             ("wflux1", st.wellSol[0].flux), ("wflux2", st.wellSol[1].flux),
             ("bhp1", st.wellSol[0].pressure), ("bhp2", st.wellSol[1].pressure))
    for key, val in pairs:
        ok, e = close(val, ref[key], 1e-9)
        assert ok, f"{key}: rel {e:.2e}"


def test_hydrostatic_column_matches_mrst():
    pa, pi, g, rock, solver = _device_case([1, 1, 30], [1.0, 1.0, 300.0], 100 * og.MILLI * og.DARCY, 0.2)
    T = solver.compute_trans(rock)
    bd = np.any(g.neighbors == -1, axis=1)
    top = np.flatnonzero(bd & (g.face_centroids[:, 2] < 1e-10))
    st = solver.solve(pi.init_res_sol(g, 0.0), T, pi.init_single_fluid(MU1, 1000.0),
                      bc=pi.add_bc(None, top, "pressure", 100 * og.BARSA), gravity=pi.gravity_vector(3))
    ref = golden("p2_gravity.npz")
    ok, e = close(st.pressure, ref["p"], 1e-9)
    assert ok, f"p: rel {e:.2e}"
    assert np.abs(st.flux - ref["flux"]).max() < 1e-12


def test_incomp_field_case_matches_oracle():
    """The 32x32x8 single-phase case: lognormal permeability (median 100 mD,
    sigma 1), rate injector and bhp producer columns, gravity on."""
    rng = np.random.default_rng(0)
    perm = 100 * og.MILLI * og.DARCY * np.exp(1.0 * rng.standard_normal(32 * 32 * 8))
    pa, pi, g, rock, solver = _device_case([32, 32, 8], [320.0, 320.0, 16.0], perm, 0.2)
    T = solver.compute_trans(rock)
    gv = pi.gravity_vector(3)
    inj = 1024 * np.arange(8)
    prod = 31 + 32 * 31 + 1024 * np.arange(8)
    w = pa.add_well(None, g, rock, inj, "rate", 1e-3, compi=[1.0], gravity=gv)
    w = pa.add_well(w, g, rock, prod, "bhp", 100 * og.BARSA, compi=[1.0], gravity=gv)
    st = solver.solve(pi.init_res_sol(g, 0.0), T, pi.init_single_fluid(MU1, 1000.0), wells=w, gravity=gv)
    # oracle
    go = og.cart_grid([32, 32, 8], [320.0, 320.0, 16.0])
    ro = og.make_rock(go, perm, 0.2)
    to = og.compute_trans(go, ro)
    wo = og.add_well(None, go, ro, inj, "rate", 1e-3, compi=[1.0], gravity=gv)
    wo = og.add_well(wo, go, ro, prod, "bhp", 100 * og.BARSA, compi=[1.0], gravity=gv)
    res = o_incomp(go, to, np.full((go.num_cells, 1), 1 / MU1), np.array([1000.0]), wells=wo, gravity=gv)
    for key, a, b in (("hT", T.cpu().numpy(), to), ("p", st.pressure, res.pressure), ("flux", st.flux, res.flux),
                      ("fp", st.facePressure, res.facePressure), ("bhp", st.wellSol[0].pressure, res.well_pressure[0]),
                      ("wflux", st.wellSol[1].flux, res.well_flux[1])):
        ok, e = close(a, b, 1e-9)
        assert ok, f"{key}: rel {e:.2e}"
    # mass balance: injected = produced (incompressible)
    q_in = st.wellSol[0].flux.sum()
    q_out = st.wellSol[1].flux.sum()
    assert abs(q_in + q_out) <= 1e-9 * abs(q_in)


# -- the rest of incomp_tpfa's surface (incomp_tpfa.py:55-63, 172-184) -----------
def _wells_case():
    """The wells-and-gravity case of MRST's phase-2 replay (10x10x3)."""
    pa, pi, g, rock, solver = _device_case([10, 10, 3], [1000.0, 1000.0, 30.0],
                                           np.tile(np.array([100.0, 100.0, 10.0]) * og.MILLI * og.DARCY,
                                                   (300, 1)), 0.3)
    gv = pi.gravity_vector(3)
    w = pa.add_well(None, g, rock, [0, 100, 200], "rate", 5e-3, compi=[1.0], gravity=gv)
    w = pa.add_well(w, g, rock, [99, 199, 299], "bhp", 100 * og.BARSA, compi=[1.0], gravity=gv)
    go = og.cart_grid([10, 10, 3], [1000.0, 1000.0, 30.0])
    ro = og.make_rock(go, np.tile(np.array([100.0, 100.0, 10.0]) * og.MILLI * og.DARCY, (300, 1)), 0.3)
    wo = og.add_well(None, go, ro, [0, 100, 200], "rate", 5e-3, compi=[1.0], gravity=gv)
    wo = og.add_well(wo, go, ro, [99, 199, 299], "bhp", 100 * og.BARSA, compi=[1.0], gravity=gv)
    return pa, pi, g, rock, w, go, ro, wo, gv


def test_matrix_output_and_lin_solve_match_the_oracle():
    """MatrixOutput returns the oracle's assembled system (null-space pin and
    well rows included) to 1e-13; LinSolve(A, b) on it gives the device
    solve's state; reduce returns the wells' Schur complement with the same
    pressures.  Gravity from the process-wide setting (on), as the
    reference's default."""
    import scipy.sparse.linalg as spla
    from pymrst_amd.gravity import isolated_gravity
    pa, pi, g, rock, w, go, ro, wo, gv = _wells_case()
    T = pi.compute_trans(g, rock)
    fluid = pi.init_single_fluid(MU1, 1000.0)
    ref = o_incomp(go, og.compute_trans(go, ro), np.full((300, 1), 1 / MU1), np.array([1000.0]), wells=wo,
                   gravity=gv)
    calls = []

    def direct(A, b):
        calls.append(A.shape)
        return spla.spsolve(A.tocsc(), b)

    with isolated_gravity("reset", "on"):
        dev = pi.incomp_tpfa(pi.init_res_sol(g, 0.0), g, T, fluid, W=w, MatrixOutput=True)
        host = pi.incomp_tpfa(pi.init_res_sol(g, 0.0), g, T, fluid, wells=w, LinSolve=direct)
        red = pi.incomp_tpfa(pi.init_res_sol(g, 0.0), g, T, fluid, wells=w, reduce=True, LinSolve=direct)
    Ad, Ao = dev.A.toarray(), ref.A.toarray()
    assert Ad.shape == Ao.shape == (302, 302)
    assert np.abs(Ad - Ao).max() <= 1e-13 * np.abs(Ao).max()
    assert close(dev.rhs, ref.rhs, 1e-13)[0]
    assert calls == [(302, 302), (300, 300)]
    assert red.A.shape == (300, 300) and red.rhs.size == 300
    for st in (dev, host, red):
        assert close(st.pressure, ref.pressure, 1e-9)[0]
        assert close(st.flux, ref.flux, 1e-9)[0]
        for a, b in zip(st.wellSol, ref.well_flux):
            assert close(a.flux, b, 1e-9)[0]


def test_global_gravity_default_and_use_trans():
    """gravity=None reads the process-wide setting: on equals the explicit
    vector, off equals no gravity; use_trans with the harmonic face
    transmissibilities gives the one-sided solve (incomp_tpfa.py:544-559)."""
    from pymrst_amd.gravity import isolated_gravity
    pa, pi, g, rock, w, go, ro, wo, gv = _wells_case()
    T = pi.compute_trans(g, rock)
    fluid = pi.init_single_fluid(MU1, 1000.0)
    st0 = pi.init_res_sol(g, 0.0)
    with isolated_gravity("reset", "on"):
        on = pi.incomp_tpfa(st0, g, T, fluid, wells=w)
        explicit = pi.incomp_tpfa(st0, g, T, fluid, wells=w, gravity=gv)
    with isolated_gravity("reset", "off"):
        off = pi.incomp_tpfa(st0, g, T, fluid, wells=w)
        half = pi.incomp_tpfa(st0, g, T, fluid, wells=w)
        t = np.asarray(T)
        Tf = 1.0 / np.bincount(g.cell_faces.ravel(), 1.0 / t, minlength=g.num_faces)
        face = pi.incomp_tpfa(st0, g, Tf, fluid, wells=w, use_trans=True)
    assert np.array_equal(on.pressure, explicit.pressure)
    assert not close(on.pressure, off.pressure, 1e-6)[0]
    assert close(face.pressure, half.pressure, 1e-9)[0]
    internal = np.all(g.neighbors >= 0, axis=1)
    assert np.abs(face.flux[internal] - half.flux[internal]).max() <= 1e-9 * np.abs(half.flux).max()


def test_periodic_pressure_jump_adds_to_the_driving_force():
    """A pressure jump d across one internal face of a 1-D chain between two
    Dirichlet ends adds to the driving force: the chain's flux becomes
    (p_L - p_R + d) / R with R = (p_L - p_R) / q(0) from the run without it."""
    from types import SimpleNamespace
    pa, pi, g, rock, solver = _device_case([12, 1, 1], [120.0, 10.0, 10.0], 100 * og.MILLI * og.DARCY, 0.2)
    T = pi.compute_trans(g, rock)
    fluid = pi.init_single_fluid(MU1, 1000.0)
    bd = np.any(g.neighbors == -1, axis=1)
    fc = g.face_centroids
    left, right = np.flatnonzero(bd & (fc[:, 0] < 1e-10)), np.flatnonzero(bd & (fc[:, 0] > 120 - 1e-10))
    bc = pi.add_bc(pi.add_bc(None, left, "pressure", 200 * og.BARSA), right, "pressure", 100 * og.BARSA)
    xfaces = np.flatnonzero(np.all(g.neighbors >= 0, axis=1) & (np.abs(fc[:, 1] - 5.0) < 1e-9)
                            & (np.abs(fc[:, 2] - 5.0) < 1e-9))
    jump_face = xfaces[5]
    d = 30 * og.BARSA
    st0 = pi.init_res_sol(g, 0.0)
    q0 = pi.incomp_tpfa(st0, g, T, fluid, bc=bc, gravity=np.zeros(3)).flux[jump_face]
    q1 = pi.incomp_tpfa(st0, g, T, fluid, bc=bc, gravity=np.zeros(3),
                        bcp=SimpleNamespace(face=np.array([jump_face]), value=np.array([d]))).flux[jump_face]
    R = 100 * og.BARSA / q0
    assert q1 == pytest.approx((100 * og.BARSA + d) / R, rel=1e-9)
