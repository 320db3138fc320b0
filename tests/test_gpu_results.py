"""Restart on the device: a schedule run into an on-disk result store,
stopped after two report steps and resumed by a fresh model and a fresh
handler from the files alone (the oil-water model's run memory, its previous
well cdp, restored from the report store), matches the uninterrupted run,
which matches the oracle's
(same Newton counts; states within 1e-8, the parity bar: the fresh model's
pressure AMG aggregates from the resumed step's matrix, so solves differ in
rounding);
device states go into the store as their host states.  Oil-water and
black-oil."""
import numpy as np
import pytest

import oracle.grid as og

from cases import build_device, build_oracle, small_gravity_inputs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _handlers(root):
    from pymrst_amd.results import ResultHandler
    return (ResultHandler(root, write_to_disk=True, store_in_memory=False),
            ResultHandler(root, data_prefix="report", write_to_disk=True, store_in_memory=False))


def test_oil_water_restart_from_disk(tmp_path):
    from oracle.ow import simulate_schedule as oracle_schedule
    from pymrst_amd.ow import CPRSolver, OWState, simulate_schedule_ad
    from pymrst_amd.results import simulate_with_output
    inp = small_gravity_inputs()
    dts = np.concatenate([inp["dts"], [20 * og.DAY]])
    md, s0 = build_device(inp, CPRSolver(rtol=1e-10))
    full, full_its = simulate_schedule_ad(s0, md, dts)
    # the uninterrupted run is the oracle's (parity), the resumed one is the uninterrupted one
    mo, so = build_oracle(inp)
    ref, ref_its = oracle_schedule(so, mo, dts)
    assert full_its.tolist() == ref_its.tolist()
    st, rep = _handlers(tmp_path)
    simulate_with_output(s0, md, dts[:2], st, rep)
    del md
    md2, s0b = build_device(inp, CPRSolver(rtol=1e-10))
    st, rep = _handlers(tmp_path)
    states, its = simulate_with_output(s0b, md2, dts, st, rep)
    assert its[2:].tolist() == full_its[2:].tolist() and its[:2].tolist() == [0, 0]
    assert [r["iterations"] for r in rep][2:] == full_its[2:].tolist()
    for a, b, o in zip(states, full, ref):
        assert isinstance(a, OWState)
        assert np.abs(a.pressure - b.pressure).max() <= 1e-8 * np.abs(b.pressure).max()
        assert np.abs(a.sW - b.sW).max() <= 1e-8
        assert np.abs(a.pressure - o.pressure).max() <= 1e-8 * np.abs(o.pressure).max()
        assert np.abs(a.sW - o.sW).max() <= 1e-8
    # a device state goes in as its host state
    dev = md2.device_state(full[-1])
    st[9] = dev
    back = st[9]
    assert isinstance(back, OWState) and back.pressure.tobytes() == full[-1].pressure.tobytes()


def test_black_oil_restart_from_disk(tmp_path):
    import pymrst_amd as pa
    from oracle.bo import synthetic_deck
    from pymrst_amd.bo import BOState, ThreePhaseBlackOilModelAD, init_deck_blackoil_fluid
    from pymrst_amd.ow import CPRSolver, simulate_schedule_ad
    from pymrst_amd.results import simulate_with_output
    dims = (6, 5, 3)
    nx, ny, nz = dims
    g = pa.cart_grid(list(dims), [30.0 * nx, 30.0 * ny, 5.0 * nz])
    rng = np.random.default_rng(7)
    rock = pa.make_rock(g, 100 * og.MILLI * og.DARCY * np.exp(0.5 * rng.standard_normal(g.num_cells)),
                        0.2 + 0.05 * rng.random(g.num_cells))
    nc = g.num_cells
    ws = pa.add_well(None, g, rock, [0], "rate", 5e-4, compi=np.array([0.0, 0.0, 1.0]))
    ws = pa.add_well(ws, g, rock, [nc - 1 - nx * ny, nc - 1], "bhp", 150 * og.BARSA,
                     compi=np.array([0.0, 1.0, 0.0]))

    def model():
        return ThreePhaseBlackOilModelAD(g, rock, init_deck_blackoil_fluid(synthetic_deck()), ws,
                                         gravity=og.NORM_GRAVITY, linear_solver=CPRSolver(rtol=1e-12))

    md = model()
    s0 = md.initial_state(200 * og.BARSA, 0.25, 0.0, rs0=45.0)
    s0.bhp[:] = [200 * og.BARSA, 150 * og.BARSA]
    dts = np.array([1.0, 4.0, 10.0, 20.0]) * og.DAY
    full, full_its = simulate_schedule_ad(s0, md, dts)
    # the uninterrupted run against the oracle's
    from oracle.bo import BOFluid, BOModel, simulate_schedule as oracle_schedule
    go = og.cart_grid(list(dims), [30.0 * nx, 30.0 * ny, 5.0 * nz])
    ro = og.make_rock(go, rock.perm, rock.poro)
    wo = og.add_well(None, go, ro, [0], "rate", 5e-4, compi=np.array([0.0, 0.0, 1.0]))
    wo = og.add_well(wo, go, ro, [nc - 1 - nx * ny, nc - 1], "bhp", 150 * og.BARSA, compi=np.array([0.0, 1.0, 0.0]))
    N, T = og.internal_connections(go, og.compute_trans(go, ro))[:2]
    mo = BOModel(go, ro, BOFluid(synthetic_deck()), wo, N, T, gravity=og.NORM_GRAVITY)
    so = mo.initial_state(200 * og.BARSA, 0.25, 0.0, rs0=45.0)
    so.bhp[:] = [200 * og.BARSA, 150 * og.BARSA]
    so.qWs[:], so.qOs[:], so.qGs[:] = s0.qWs, s0.qOs, s0.qGs
    ref, ref_its = oracle_schedule(so, mo, dts)
    assert full_its.tolist() == ref_its.tolist()
    for a, o in zip(full, ref):
        np.testing.assert_array_equal(a.status, o.status)
        np.testing.assert_allclose(a.pressure, o.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.s, o.s, rtol=0, atol=1e-8)
    st, rep = _handlers(tmp_path)
    simulate_with_output(s0, md, dts[:2], st, rep)
    del md
    st, rep = _handlers(tmp_path)
    states, its = simulate_with_output(s0, model(), dts, st, rep)
    assert its[2:].tolist() == full_its[2:].tolist()
    for a, b in zip(states, full):
        assert isinstance(a, BOState)
        np.testing.assert_array_equal(a.status, b.status)
        np.testing.assert_allclose(a.pressure, b.pressure, rtol=1e-8)
        np.testing.assert_allclose(a.s, b.s, rtol=0, atol=1e-8)
        np.testing.assert_allclose(a.rs, b.rs, rtol=1e-8, atol=1e-8)
