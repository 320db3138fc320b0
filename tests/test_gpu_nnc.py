"""General connection lists on the device: a faulted grid whose connection
list drops the regular x-connections across part of a fault and adds
non-neighbouring pairs (displaced columns, one long-range pair listed in
reversed order), as corner-point grids with NNCs present them
(pymrst/modules/ad_core/utils/setup_operators_tpfa.py:129-137 takes N and T
as given).  The device models take the same N and T as the oracle: residual,
pattern and J x at a perturbed state, and whole schedules (same Newton
counts, states within 1e-8) for oil-water and black-oil.  A pair listed twice
is an error."""
import ctypes

import numpy as np
import pytest

import oracle.grid as og
from oracle.ow import simulate_schedule

from cases import build_device, build_oracle, device_to_reference_perm, faulted_connections, small_gravity_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

DIMS = dict(nx=8, ny=6, nz=4)


def test_faulted_assembly_matches_oracle():
    from pymrst_amd import _lib
    from pymrst_amd.ow import DeviceOWState
    inp = small_gravity_inputs(**DIMS)
    mo, so = build_oracle(inp, connections=faulted_connections)
    md, _ = build_device(inp, connections=faulted_connections)
    nc, nw = mo.nc, len(mo.wells)
    st = so.copy()
    rng = np.random.default_rng(5)
    st.pressure = st.pressure + rng.normal(0, 2e5, nc)
    st.sW = np.clip(st.sW + rng.normal(0, 0.1, nc), 0.0, 1.0)
    dt = 10 * og.DAY
    cdp = mo.compute_cdp_per_well(so)
    r_ref, J_ref, _ = mo.residual_and_jacobian(so, st, dt, cdp)
    dev = torch.device("cuda")
    md.assemble_device(DeviceOWState.from_host(so, dev), DeviceOWState.from_host(st, dev), dt,
                       torch.as_tensor(np.concatenate(cdp), device=dev))
    perm = device_to_reference_perm(nc, nw)
    assert np.abs(md.residual_device().cpu().numpy() - r_ref[perm]).max() <= 1e-12 * np.abs(r_ref).max()
    nnzb = ctypes.c_int32()
    _lib.check(md.lib.pmx_ow_nnzb(md.h, ctypes.byref(nnzb)))
    rowptr = np.zeros(nc + 1, np.int32)
    col = np.zeros(nnzb.value, np.int32)
    _lib.check(md.lib.pmx_ow_export_pattern(md.h, _lib.ptr(rowptr), _lib.ptr(col)))
    N = mo.ops.N
    expect = set(zip(N[:, 0], N[:, 1])) | set(zip(N[:, 1], N[:, 0])) | {(i, i) for i in range(nc)}
    got = {(i, int(c)) for i in range(nc) for c in col[rowptr[i]:rowptr[i + 1]]}
    assert got == expect
    assert (nc - 1, 0) in got and (0, nc - 1) in got          # the long-range pair, both blocks
    nx, ny = DIMS["nx"], DIMS["ny"]
    a = nx // 2 - 1                                             # (i0-1, 0, 0) -> (i0, 0, 1)
    assert (a, a + 1 + nx * ny) in got and (a, a + 1) not in got
    x_ref = rng.normal(size=J_ref.shape[0])
    y_ref = J_ref @ x_ref
    x_dev = torch.as_tensor(x_ref[perm], device=dev)
    y_dev = torch.empty_like(x_dev)
    _lib.check(md.lib.pmx_ow_matvec(md.h, _lib.ptr(x_dev), _lib.ptr(y_dev)))
    ys = np.asarray(abs(J_ref) @ np.abs(x_ref)).ravel()
    assert np.all(np.abs(y_dev.cpu().numpy() - y_ref[perm]) <= 1e-12 * (ys[perm] + 1e-300))


@pytest.mark.parametrize("prec", [64, 32])
def test_faulted_oil_water_schedule_matches_oracle(prec):
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    inp = small_gravity_inputs(**DIMS)
    mo, so = build_oracle(inp, connections=faulted_connections)
    ostates, oits = simulate_schedule(so, mo, inp["dts"])
    md, sd = build_device(inp, CPRSolver(rtol=1e-10, precision=prec), connections=faulted_connections)
    states, its = simulate_schedule_ad(sd, md, inp["dts"])
    assert np.array_equal(its, oits), f"gpu {its.tolist()} oracle {oits.tolist()}"
    p, po = np.vstack([s.pressure for s in states]), np.vstack([s.pressure for s in ostates])
    assert np.abs(p - po).max() <= 1e-8 * np.abs(po).max()
    sw, swo = np.vstack([s.sW for s in states]), np.vstack([s.sW for s in ostates])
    assert np.abs(sw - swo).max() <= 1e-8
    # not vacuous: the fault moves the answer far beyond the tolerance
    m0, s0 = build_oracle(inp)
    plain, _ = simulate_schedule(s0, m0, inp["dts"])
    assert np.abs(plain[-1].pressure - ostates[-1].pressure).max() > 1e-3 * np.abs(po).max()


def test_faulted_black_oil_schedule_matches_oracle():
    import pymrst_amd as pa
    from oracle.bo import BOFluid, BOModel, simulate_schedule as bo_schedule, synthetic_deck
    from pymrst_amd.bo import ThreePhaseBlackOilModelAD, init_deck_blackoil_fluid
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    dims = (6, 5, 3)
    nx, ny, nz = dims
    phys = [30.0 * nx, 30.0 * ny, 5.0 * nz]
    g = og.cart_grid(list(dims), phys)
    rng = np.random.default_rng(7)
    perm = 100 * og.MILLI * og.DARCY * np.exp(0.5 * rng.standard_normal(g.num_cells))
    poro = 0.2 + 0.05 * rng.random(g.num_cells)
    rock = og.make_rock(g, perm, poro)
    N, T = faulted_connections(g, *og.internal_connections(g, og.compute_trans(g, rock))[:2])
    nc = g.num_cells
    ws = og.add_well(None, g, rock, [0], "rate", 5e-4, compi=[0.0, 0.0, 1.0])
    ws = og.add_well(ws, g, rock, [nc - 1 - nx * ny, nc - 1], "bhp", 150 * og.BARSA, compi=[0.0, 1.0, 0.0])
    deck = synthetic_deck()
    mo = BOModel(g, rock, BOFluid(deck), ws, N, T, gravity=og.NORM_GRAVITY)
    gd = pa.cart_grid(list(dims), phys)
    wd = [pa.Well(cells=np.asarray(w.cells), type=w.type, val=w.val, WI=np.asarray(w.WI, float),
                  dZ=np.zeros(len(w.cells)), compi=np.asarray(w.compi, float)) for w in ws]
    md = ThreePhaseBlackOilModelAD(gd, pa.make_rock(gd, perm, poro), init_deck_blackoil_fluid(deck), wd,
                                   gravity=og.NORM_GRAVITY, linear_solver=CPRSolver(rtol=1e-12), T=T, N=N)
    st0 = mo.initial_state(200 * og.BARSA, 0.25, 0.0, rs0=45.0)
    st0.bhp[:] = [200 * og.BARSA, 150 * og.BARSA]
    dts = np.array([1.0, 4.0, 10.0, 20.0]) * og.DAY
    ref_states, ref_its = bo_schedule(st0, mo, dts)
    states, its = simulate_schedule_ad(st0, md, dts)
    np.testing.assert_array_equal(its, ref_its)
    for a, b in zip(states, ref_states):
        np.testing.assert_array_equal(a.status, b.status)
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.s, b.s, rtol=0, atol=1e-8)
        np.testing.assert_allclose(a.rs, b.rs, rtol=1e-8, atol=1e-8)


def test_a_pair_listed_twice_is_an_error():
    from pymrst_amd import _lib
    inp = small_gravity_inputs(**DIMS)

    def doubled(g, N, T):
        return np.vstack([N, N[:1]]), np.concatenate([T, T[:1]])

    with pytest.raises(_lib.PmxError, match="cell pair twice"):
        build_device(inp, connections=doubled)
