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
