"""Build the oracle's oil-water model from a workload description (any
object with the fields of pymrst_amd.workloads.OWCase) -- test
infrastructure only."""
import numpy as np

from . import grid as og
from .ow import OWFluid, OWModel


def model_from_case(c, linear_solver=None):
    g = og.cart_grid(c.dims, c.phys)
    rock = og.make_rock(g, c.perm, c.poro)
    grav = np.array([0.0, 0.0, c.gravity]) if c.gravity else None
    wells = None
    for cells, kind, val, compi in c.wells:
        wells = og.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi),
                            sign=1 if kind == "rate" else -1, gravity=grav)
    fl = OWFluid(mu=np.array(c.mu, float), rhoS=np.array(c.rho, float), n=np.array(c.n, float),
                 c=np.array(c.c, float), p_ref=c.p_ref, c_r=c.c_r)
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    m = OWModel(g, rock, fl, wells, N, T, gravity=c.gravity, linear_solver=linear_solver)
    return m, m.initial_state(c.p0, c.sw0)


def bo_model_from_case(c):
    """The oracle's black-oil model and initial state from a workload with
    the fields of pymrst_amd.workloads.BOCase."""
    from .bo import BOFluid, BOModel
    g = og.cart_grid(c.dims, c.phys)
    rock = og.make_rock(g, c.perm, c.poro)
    grav = np.array([0.0, 0.0, c.gravity]) if c.gravity else None
    wells = None
    for cells, kind, val, compi in c.wells:
        wells = og.add_well(wells, g, rock, cells, kind, val, compi=np.array(compi),
                            sign=1 if kind == "rate" else -1, gravity=grav)
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    m = BOModel(g, rock, BOFluid(c.deck), wells, N, T, gravity=c.gravity)
    return m, m.initial_state(c.p_init, c.sw0, c.sg_init, rs0=c.rs0)
