"""Synthetic workloads for the oil-water Newton benchmark.

``field_waterflood`` is the fully-implicit oil-water field case: a
60 x 220 x 85 Cartesian grid (1,122,000 cells of 6.1 x 3.05 x 0.61 m),
layered lognormal permeability (per-layer median log-uniform on 1-1000 mD,
in-layer sigma(ln k) = 1.5, kz/kx = 0.1), porosity 0.05 + 0.25 Phi(z)
correlated with ln k, Corey 2/2 relative permeabilities, mu_w = 0.3 cP,
mu_o = 3 cP, c_w = c_o = 1e-9 / Pa, c_r = 1e-10 / Pa, a five-spot (one
centre water injector at 5000 bbl/day -- the SPE10 Model 2 rate for this
SPE10 Model 2 geometry -- and four corner producers at 150 bar), initial
pressure 200 bar and water saturation 0.2, gravity on, 10 steps of 30 days.  Everything is drawn from one seeded generator, so the same inputs
can be rebuilt anywhere.
"""
from dataclasses import dataclass
from typing import Optional

import numpy as np
from scipy.special import ndtr

from .grid import BARSA, CENTI, DARCY, DAY, MILLI, NORM_GRAVITY, POISE


@dataclass
class OWCase:
    dims: tuple
    phys: tuple
    perm: np.ndarray        # (nc, 3) SI
    poro: np.ndarray
    mu: tuple
    rho: tuple
    n: tuple
    c: tuple
    p_ref: float
    c_r: float
    wells: list             # (cells, type, val, compi)
    p0: float
    sw0: float
    dts: np.ndarray
    gravity: float
    p_init: Optional[np.ndarray] = None   # per-cell initial pressure (None: p0 everywhere)

    @property
    def num_cells(self):
        return int(np.prod(self.dims))


STB = 42.0 * 231.0 * 0.0254 ** 3    # stock-tank barrel in m^3


def field_waterflood(dims=(60, 220, 85), cell=(6.1, 3.05, 0.61), seed=1, steps=10, step_days=30.0,
                     inj_rate=5000.0 * STB / DAY, prod_bhp=150 * BARSA, sigma=1.5) -> OWCase:
    nx, ny, nz = dims
    nc = nx * ny * nz
    rng = np.random.default_rng(seed)
    medians_md = 10.0 ** rng.uniform(0.0, 3.0, nz)
    z = rng.standard_normal(nc)
    layer = np.arange(nc) // (nx * ny)
    kx = np.exp(np.log(medians_md[layer]) + sigma * z) * MILLI * DARCY
    perm = np.column_stack([kx, kx, 0.1 * kx])
    poro = 0.05 + 0.25 * ndtr(z)
    column = lambda i, j: i + nx * j + nx * ny * np.arange(nz)
    wells = [(column(nx // 2, ny // 2), "rate", inj_rate, (1.0, 0.0))]
    for i, j in ((0, 0), (nx - 1, 0), (0, ny - 1), (nx - 1, ny - 1)):
        wells.append((column(i, j), "bhp", prod_bhp, (0.0, 1.0)))
    return OWCase(dims=dims, phys=(nx * cell[0], ny * cell[1], nz * cell[2]), perm=perm, poro=poro,
                  mu=(0.3 * CENTI * POISE, 3.0 * CENTI * POISE), rho=(1000.0, 850.0), n=(2.0, 2.0),
                  c=(1e-9, 1e-9), p_ref=200 * BARSA, c_r=1e-10, wells=wells, p0=200 * BARSA, sw0=0.2,
                  dts=np.full(steps, step_days * DAY), gravity=NORM_GRAVITY)


def build_model(case: OWCase, linear_solver=None, device=None):
    """The device model for a case (grid, rock, wells, fluid)."""
    from . import grid as G
    from . import ow
    g = G.cart_grid(case.dims, case.phys)
    rock = G.make_rock(g, case.perm, case.poro)
    grav = np.array([0.0, 0.0, case.gravity]) if case.gravity else None
    wells = None
    for cells, kind, val, compi in case.wells:
        wells = G.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi),
                           sign=1 if kind == "rate" else -1, gravity=grav)
    fluid = ow.init_simple_adi_fluid(case.mu, case.rho, case.n, case.c, case.p_ref, c_r=case.c_r)
    model = ow.TwoPhaseOilWaterModelAD(g, rock, fluid, wells, gravity=case.gravity,
                                       linear_solver=linear_solver, device=device)
    st = model.initial_state(case.p0, case.sw0)
    if case.p_init is not None:
        st.pressure = np.asarray(case.p_init, float).copy()
    return model, st


def adjoint_problem(dims=(64, 64, 16), cell=(6.1, 3.05, 0.61), seed=4, steps=20, step_days=15.0,
                    inj_rate=2.5e-3, prod_bhp=150 * BARSA, sigma=1.0, solver=None, checkpoint_every=1):
    """The log-permeability adjoint case: oil-water with the field case's
    fluids on a `dims` grid, ln k = ln(100 mD) + sigma z (z seeded standard
    normal, kz/kx 0.1), porosity 0.2, a centre rate injector (2.5e-3 m^3/s,
    about half a pore volume over the schedule at 64x64x16) and four corner
    producers at 150 bar, no gravity, `steps` steps of `step_days` days.
    Objective: sum over report steps of the producers' squared water-rate
    mismatch against seeded targets (one quarter of the injection rate, 20 %
    seeded spread).  Returns (problem, ln k)."""
    from . import grid as G
    from . import ow
    from .adjoint import AdjointProblem
    nx, ny, nz = dims
    nc = nx * ny * nz
    rng = np.random.default_rng(seed)
    logk = np.log(100 * MILLI * DARCY) + sigma * rng.standard_normal(nc)
    g = G.cart_grid(dims, (nx * cell[0], ny * cell[1], nz * cell[2]))
    column = lambda i, j: i + nx * j + nx * ny * np.arange(nz)
    specs = [dict(cells=column(nx // 2, ny // 2), type="rate", val=inj_rate, compi=[1.0, 0.0], sign=1)]
    for i, j in ((0, 0), (nx - 1, 0), (0, ny - 1), (nx - 1, ny - 1)):
        specs.append(dict(cells=column(i, j), type="bhp", val=prod_bhp, compi=[0.0, 1.0], sign=-1))
    targets = -0.25 * inj_rate * (1.0 + 0.2 * rng.standard_normal((steps, 4)))
    fluid = ow.init_simple_adi_fluid((0.3 * CENTI * POISE, 3.0 * CENTI * POISE), (1000.0, 850.0), (2.0, 2.0),
                                     (1e-9, 1e-9), 200 * BARSA, c_r=1e-10)
    prob = AdjointProblem(g, np.full(nc, 0.2), fluid, specs, np.full(steps, step_days * DAY), 200 * BARSA, 0.2,
                          solver=solver, targets=targets, checkpoint_every=checkpoint_every)
    return prob, logk


def ensemble_cases(members=256, seed=3, steps=5, step_days=30.0, rate_scale=0.25):
    """The uneven ensemble: member m has nx, ny ~ U{20..64}, nz ~ U{5..16},
    sigma(ln k) ~ U(0.5, 2.0) and its own seed, the field case's fluids,
    cells and five-spot, with the injection rate scaled by the member's
    cell count against the 60x220x85 field times `rate_scale` (0.25: a
    quarter of the field's injection per cell, which every member alone
    takes in 30-day steps without cuts), `steps` steps of `step_days`
    days."""
    rng = np.random.default_rng(seed)
    cell_dz = 0.61
    cases = []
    for m in range(members):
        nx, ny = (int(v) for v in rng.integers(20, 65, 2))
        nz = int(rng.integers(5, 17))
        sigma = float(rng.uniform(0.5, 2.0))
        rate = rate_scale * 5000.0 * STB / DAY * (nx * ny * nz) / (60 * 220 * 85)
        c = field_waterflood(dims=(nx, ny, nz), seed=int(rng.integers(1 << 31)), steps=steps,
                             step_days=step_days, inj_rate=rate, sigma=sigma)
        # start in hydrostatic equilibrium of the initial fluid column
        rho0 = c.sw0 * c.rho[0] + (1.0 - c.sw0) * c.rho[1]
        z = (np.arange(c.num_cells) // (nx * ny) + 0.5) * cell_dz
        c.p_init = c.p0 + c.gravity * rho0 * (z - 0.5 * cell_dz)
        cases.append(c)
    return cases


@dataclass
class BOCase:
    dims: tuple
    phys: tuple
    perm: np.ndarray
    poro: np.ndarray
    deck: dict
    wells: list             # (cells, type, val, compi)
    p_init: np.ndarray
    sw0: float
    sg_init: np.ndarray
    rs0: Optional[float]    # None: oil saturated at the initial pressure
    dts: np.ndarray
    gravity: float

    @property
    def num_cells(self):
        return int(np.prod(self.dims))


def synthetic_bo_deck(nrows: int = 20, p_lo=50 * BARSA, p_hi=400 * BARSA) -> dict:
    """The synthetic monotone black-oil PVT: `nrows` saturated PVTO rows over
    50-400 bar with Bo 1.05 -> 1.35, Rs 0 -> 150 Sm3/Sm3, mu_o 2 -> 0.8 cP,
    each with one undersaturated row 50 bar above (Bo -1 %, mu_o +5 %); PVDG
    with Bg from a real-gas Z factor 0.95 -> 0.85 at 373.15 K and mu_g
    0.012 -> 0.024 cP; Corey-2 SWOF/SGOF with connate water 0.2; PVTW
    (B 1.01 at 200 bar, c 4e-10 /Pa, 0.5 cP), ROCK c 5e-10 /Pa at 200 bar;
    surface densities 1000 / 850 / 0.9 kg/m3."""
    p = np.linspace(p_lo, p_hi, nrows)
    t = (p - p_lo) / (p_hi - p_lo)
    rs = 150.0 * t
    bo = 1.05 + 0.30 * t
    muo = (2.0 - 1.2 * t) * CENTI * POISE
    pvto = [(rs[k], [[p[k], bo[k], muo[k]], [p[k] + 50 * BARSA, bo[k] * 0.99, muo[k] * 1.05]])
            for k in range(nrows)]
    Z = 0.95 - 0.10 * t
    Tk, p_sc, T_sc = 373.15, 1.01325e5, 288.71
    Bg = Z * Tk * p_sc / (p * T_sc)
    mug = (0.012 + 0.012 * t) * CENTI * POISE
    sw = np.linspace(0.2, 1.0, 9)
    swof = np.column_stack([sw, ((sw - 0.2) / 0.8) ** 2, ((1.0 - sw) / 0.8) ** 2, np.zeros_like(sw)])
    sg = np.linspace(0.0, 0.8, 9)
    sgof = np.column_stack([sg, (sg / 0.8) ** 2, ((0.8 - sg) / 0.8) ** 2, np.zeros_like(sg)])
    return dict(rhoS=[1000.0, 850.0, 0.9],
                pvtw=dict(p_ref=200 * BARSA, b_ref=1.01, c=4e-10, mu_ref=0.5 * CENTI * POISE, viscosibility=0.0),
                rock_cp=dict(p_ref=200 * BARSA, c=5e-10), pvdg=np.column_stack([p, Bg, mug]), pvto=pvto,
                swof=swof, sgof=sgof)


def blackoil_field(dims=(100, 100, 20), cell=(20.0, 20.0, 2.0), seed=2, steps=20, step_days=15.0,
                   gas_rate=1.0e6 / DAY, prod_bhp=150 * BARSA, p_top=250 * BARSA, cap_layers=4, sg_cap=0.25,
                   sigma=1.5) -> BOCase:
    """The three-phase live-oil case: a `dims` Cartesian grid (200,000 cells
    of 20 x 20 x 2 m by default), layered lognormal permeability as in the
    oil-water field case (per-layer median log-uniform on 1-1000 mD,
    in-layer sigma(ln k) = `sigma`, kz/kx = 0.1, porosity 0.05 + 0.25 Phi(z)),
    the synthetic PVT deck and connate water.  A gas cap (gas saturation
    `sg_cap`) fills the top `cap_layers` layers over oil saturated at its
    pressure, in hydrostatic equilibrium from `p_top` at the top.  A centre
    gas injector (1e6 Sm3/day) is completed in the gas cap, four corner
    producers at 150 bar in the oil zone below the cap's first layer of
    oil; gravity on; `steps` steps of `step_days` days.  Drawdown below the
    bubble point liberates gas around the producers, so cells switch from
    Rs to Sg as primary variable during the run."""
    nx, ny, nz = dims
    nc = nx * ny * nz
    rng = np.random.default_rng(seed)
    medians_md = 10.0 ** rng.uniform(0.0, 3.0, nz)
    z = rng.standard_normal(nc)
    layer = np.arange(nc) // (nx * ny)
    kx = np.exp(np.log(medians_md[layer]) + sigma * z) * MILLI * DARCY
    perm = np.column_stack([kx, kx, 0.1 * kx])
    poro = 0.05 + 0.25 * ndtr(z)
    column = lambda i, j, ks: i + nx * j + nx * ny * np.asarray(ks)
    cap = np.arange(cap_layers)
    oil = np.arange(cap_layers + 1, nz)
    wells = [(column(nx // 2, ny // 2, cap), "rate", gas_rate, (0.0, 0.0, 1.0))]
    for i, j in ((0, 0), (nx - 1, 0), (0, ny - 1), (nx - 1, ny - 1)):
        wells.append((column(i, j, oil), "bhp", prod_bhp, (0.0, 1.0, 0.0)))
    sw0 = 0.2
    sg_layer = np.where(np.arange(nz) < cap_layers, sg_cap, 0.0)
    # hydrostatic column with the mixture density of each layer (oil about
    # 750, gas about 190 kg/m3 at reservoir conditions)
    rho = sw0 * 1000.0 + sg_layer * 190.0 + (1.0 - sw0 - sg_layer) * 750.0
    p_layer = np.empty(nz)
    p_layer[0] = p_top
    for k in range(1, nz):
        p_layer[k] = p_layer[k - 1] + NORM_GRAVITY * 0.5 * (rho[k - 1] + rho[k]) * cell[2]
    return BOCase(dims=dims, phys=(nx * cell[0], ny * cell[1], nz * cell[2]), perm=perm, poro=poro,
                  deck=synthetic_bo_deck(), wells=wells, p_init=p_layer[layer], sw0=sw0, sg_init=sg_layer[layer],
                  rs0=None, dts=np.full(steps, step_days * DAY), gravity=NORM_GRAVITY)


def build_bo_model(case: BOCase, linear_solver=None, device=None):
    """The device black-oil model and initial state for a case."""
    from . import grid as G
    from .bo import ThreePhaseBlackOilModelAD, init_deck_blackoil_fluid
    g = G.cart_grid(case.dims, case.phys)
    rock = G.make_rock(g, case.perm, case.poro)
    grav = np.array([0.0, 0.0, case.gravity]) if case.gravity else None
    wells = None
    for cells, kind, val, compi in case.wells:
        wells = G.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi),
                           sign=1 if kind == "rate" else -1, gravity=grav)
    model = ThreePhaseBlackOilModelAD(g, rock, init_deck_blackoil_fluid(case.deck), wells, gravity=case.gravity,
                                      linear_solver=linear_solver, device=device)
    st = model.initial_state(case.p_init, case.sw0, case.sg_init, rs0=case.rs0)
    return model, st


def incomp_case(dims=(32, 32, 8), phys=(320.0, 320.0, 16.0), seed=0):
    """The incompTPFA workload: a 32x32x8 grid of 10x10x2 m cells, isotropic
    cell-wise lognormal permeability (median 100 mD, sigma(ln k) 1.0, seeded),
    porosity 0.2, one fluid of 1 cP and 1000 kg/m3, a rate injector of
    1e-3 m3/s in column (0, 0) and a 100 bar producer in column (31, 31), all
    layers, other boundaries closed, gravity on.  Returns a dict with the
    product's grid, rock, wells, fluid and the raw inputs (for the oracle)."""
    from . import grid as G
    from . import incomp as pi
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    perm = 100 * MILLI * DARCY * np.exp(1.0 * rng.standard_normal(nx * ny * nz))
    g = G.cart_grid(dims, phys)
    rock = G.make_rock(g, perm, 0.2)
    gv = pi.gravity_vector(3)
    specs = [(nx * ny * np.arange(nz), "rate", 1e-3), (nx - 1 + nx * (ny - 1) + nx * ny * np.arange(nz), "bhp",
                                                        100 * BARSA)]
    wells = None
    for cells, kind, val in specs:
        wells = G.add_well(wells, g, rock, cells, kind, val, compi=[1.0], gravity=gv)
    mu = 1e-3
    return dict(grid=g, rock=rock, wells=wells, fluid=pi.init_single_fluid(mu, 1000.0), gravity=gv, dims=dims,
                phys=phys, perm=perm, poro=0.2, mu=mu, rho=1000.0, well_specs=specs)
