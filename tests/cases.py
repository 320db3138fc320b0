"""Test cases built twice: once with the oracle (checker) and once with the
product's own setup code, from the same closed-form inputs."""
import numpy as np

import oracle.grid as og
from oracle.ow import OWFluid, OWModel

GOLDEN = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def golden(name):
    return np.load(f"{GOLDEN}/{name}", allow_pickle=False)


# ---------------------------------------------------------------------------
def paired_inputs():
    """pymrst/modules/ad_core/examples/paired_heterogeneous_waterflood.py:58-101."""
    nx, ny, nz = 15, 15, 3
    cell = np.arange(nx * ny * nz)
    i, j, k = cell % nx, (cell // nx) % ny, cell // (nx * ny)
    lm = np.array([1.0, 0.35, 1.8])[k]
    ch = 1.0 + 0.55 * np.sin((i + 1) * 0.63) * np.cos((j + 1) * 0.41)
    pmd = np.clip(120.0 * lm * ch, 15.0, None)
    poro = np.clip(0.16 + 0.055 * (pmd / pmd.max()), 0.16, 0.24)
    col = lambda a, b: a + nx * b + nx * ny * np.arange(nz)
    return dict(dims=(nx, ny, nz), phys=(750.0, 750.0, 30.0), perm=pmd * og.MILLI * og.DARCY, poro=poro,
                mu=np.array([1.0, 4.0]) * og.CENTI * og.POISE, rho=[1000.0, 780.0], n=[2.0, 2.0],
                c=[1e-6 / og.BARSA, 8e-5 / og.BARSA], p_ref=250 * og.BARSA,
                wells=[(col(1, 1), "rate", 180.0 / og.DAY, [1.0, 0.0], 1),
                       (col(nx - 2, ny - 2), "bhp", 180 * og.BARSA, [0.0, 1.0], -1)],
                p0=250 * og.BARSA, sw0=0.15, dts=np.full(24, 30 * og.DAY), gravity=0.0, c_r=None)


def p5_inputs():
    """examples/validation/validate_phase5.py:23-37."""
    nc = 225
    return dict(dims=(15, 15, 1), phys=(300.0, 300.0, 10.0),
                perm=(50 + (11 * np.arange(nc)) % 200) * og.MILLI * og.DARCY, poro=0.25,
                mu=np.array([1.0, 5.0]) * og.CENTI * og.POISE, rho=[1000.0, 800.0], n=[2.0, 2.0],
                c=[1e-6 / og.BARSA, 1e-4 / og.BARSA], p_ref=200 * og.BARSA,
                wells=[(np.array([0]), "rate", 2e-3, [1.0, 0.0], 1),
                       (np.array([nc - 1]), "bhp", 150 * og.BARSA, [0.0, 1.0], -1)],
                p0=200 * og.BARSA, sw0=0.2, dts=np.full(8, 30 * og.DAY), gravity=0.0, c_r=None)


def small_gravity_inputs(seed=0, nx=6, ny=5, nz=4):
    """A small heterogeneous case with gravity, rock compressibility and a
    well that both injects and produces (cross flow)."""
    rng = np.random.default_rng(seed)
    nc = nx * ny * nz
    kx = np.exp(rng.normal(np.log(100.0), 1.0, nc)) * og.MILLI * og.DARCY
    perm = np.column_stack([kx, kx * 0.5, kx * 0.1])
    poro = 0.1 + 0.2 * rng.random(nc)
    col = lambda a, b: a + nx * b + nx * ny * np.arange(nz)
    return dict(dims=(nx, ny, nz), phys=(60.0, 50.0, 40.0), perm=perm, poro=poro,
                mu=np.array([0.3, 3.0]) * og.CENTI * og.POISE, rho=[1000.0, 850.0], n=[2.0, 2.0],
                c=[1e-9, 1e-9], p_ref=200 * og.BARSA,
                wells=[(col(0, 0), "rate", 2e-3, [1.0, 0.0], 1),
                       (col(nx - 1, ny - 1), "bhp", 150 * og.BARSA, [0.0, 1.0], -1),
                       (col(nx // 2, 0), "bhp", 198.5 * og.BARSA, [0.0, 1.0], -1)],
                p0=200 * og.BARSA, sw0=0.2, dts=np.full(3, 10 * og.DAY), gravity=og.NORM_GRAVITY,
                c_r=1e-10)


# ---------------------------------------------------------------------------
def faulted_connections(g, N, T, frac=0.5):
    """A Cartesian grid's connection list with a normal fault at x = nx/2,
    as a corner-point grid with displaced columns presents it: for j < ny/2
    the x-connections across the fault are dropped and cell (i0-1, j, k)
    connects to (i0, j, k+1) instead, through `frac` of the dropped
    transmissibility (non-neighbouring pairs, appended after the regular
    connections as MRST appends NNCs); one long-range pair, listed as
    (last cell, first cell), models a weak leak between the corners."""
    nx, ny, nz = g.dims
    i0 = nx // 2
    N = np.asarray(N)
    T = np.asarray(T, float)
    a = N[:, 0]
    i, j, k = a % nx, (a // nx) % ny, a // (nx * ny)
    cut = (N[:, 1] - a == 1) & (i == i0 - 1) & (j < ny // 2)
    keep = ~cut
    nnc = cut & (k < nz - 1)
    extra = np.column_stack([a[nnc], a[nnc] + 1 + nx * ny])
    t_extra = frac * T[nnc]
    leak = np.array([[g.num_cells - 1, 0]])
    N2 = np.vstack([N[keep], extra, leak]).astype(N.dtype)
    T2 = np.concatenate([T[keep], t_extra, [1e-3 * np.median(T)]])
    return N2, T2


def build_oracle(inp, linear_solver=None, connections=None):
    g = og.cart_grid(inp["dims"], inp["phys"])
    rock = og.make_rock(g, inp["perm"], inp["poro"])
    grav = np.array([0, 0, inp["gravity"]]) if inp["gravity"] else None
    wells = None
    for cells, kind, val, compi, sign in inp["wells"]:
        wells = og.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi), sign=sign, gravity=grav)
    fl = OWFluid(mu=np.asarray(inp["mu"], float), rhoS=np.asarray(inp["rho"], float),
                 n=np.asarray(inp["n"], float), c=np.asarray(inp["c"], float), p_ref=inp["p_ref"],
                 c_r=inp["c_r"])
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    if connections is not None:
        N, T = connections(g, N, T)
    m = OWModel(g, rock, fl, wells, N, T, gravity=inp["gravity"], linear_solver=linear_solver)
    return m, m.initial_state(inp["p0"], inp["sw0"])


def build_device(inp, linear_solver=None, connections=None):
    import pymrst_amd as pa
    from pymrst_amd import ow
    g = pa.cart_grid(inp["dims"], inp["phys"])
    rock = pa.make_rock(g, inp["perm"], inp["poro"])
    grav = np.array([0, 0, inp["gravity"]]) if inp["gravity"] else None
    wells = None
    for cells, kind, val, compi, sign in inp["wells"]:
        wells = pa.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi), sign=sign, gravity=grav)
    fl = ow.init_simple_adi_fluid(inp["mu"], inp["rho"], inp["n"], inp["c"], inp["p_ref"], c_r=inp["c_r"])
    N = T = None
    if connections is not None:
        go = og.cart_grid(inp["dims"], inp["phys"])
        N, T = og.internal_connections(go, og.compute_trans(go, og.make_rock(go, inp["perm"], inp["poro"])))[:2]
        N, T = connections(go, N, T)
    m = ow.TwoPhaseOilWaterModelAD(g, rock, fl, wells, gravity=inp["gravity"], linear_solver=linear_solver,
                                   N=N, T=T)
    return m, m.initial_state(inp["p0"], inp["sw0"])


def device_to_reference_perm(nc, nw):
    """Index map: device position d -> reference (variable-blocked) position."""
    d = np.arange(2 * nc + 3 * nw)
    ref = np.empty_like(d)
    c, v = d[:2 * nc] // 2, d[:2 * nc] % 2
    ref[:2 * nc] = v * nc + c
    w, k = (d[2 * nc:] - 2 * nc) // 3, (d[2 * nc:] - 2 * nc) % 3
    ref[2 * nc:] = 2 * nc + k * nw + w
    return ref
