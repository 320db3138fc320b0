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
