"""Single-phase compressible Newton oracle (test infrastructure only).

Restates run_simulation of examples/single_phase_compressible_ad.py:52-124
(MRST's flowSolverTutorialAD): a 10x10x10 grid of a 200x200x50 m box with
rock pore volume pv_r exp(c_r (p - p_ref)) and fluid density
rho_ref exp(c (p - p_ref)), TPFA fluxes with the gravity head
g avg(rho) grad(z) (the same Ops as the oil-water oracle), a horizontal
bhp-controlled producer along y (i = 1, j = 1..8, k = 4, 0-based), backward
Euler, and Newton on the combined [pressure, bhp, qs] system with a direct
solve until ||res||_2 <= 1e-5 (at most 10 iterations per step).  The
hydrostatic start is the closed form of dp/dz = g rho(p) (:45-49).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import grid as og
from .adi import ad_exp, ad_matmul, ad_sum, combine_equations, init_variables_adi, value
from .ow import Ops

NX, NY, NZ = 10, 10, 10
LX, LY, LZ = 200.0, 200.0, 50.0
PERM = 30 * og.MILLI * og.DARCY
PORO = 0.3
C_ROCK = 1e-6 / og.BARSA
P_REF = 200 * og.BARSA
MU = 5 * og.CENTI * og.POISE
C_FLUID = 1e-3 / og.BARSA
RHO_REF = 850.0
RHO_SURFACE = 750.0
BHP_TARGET = 100 * og.BARSA
NUM_STEPS = 52
TOT_TIME = 365 * og.DAY
NEWTON_TOL = 1e-5
MAX_ITS = 10


def hydrostatic_pressure(z, g=og.NORM_GRAVITY):
    """p(z) solving dp/dz = g rho_ref exp(c (p - p_ref)), p(0) = p_ref (:45-49)."""
    a = C_FLUID * g * RHO_REF * np.exp(-C_FLUID * P_REF)
    return -np.log(np.exp(-C_FLUID * P_REF) - a * z) / C_FLUID


def run_simulation(gravity=og.NORM_GRAVITY, steps=NUM_STEPS):
    """Returns dict(p_hist, bhp_hist, qs_hist, p_init, WI, dZ, nits, pv_r)."""
    g = og.cart_grid([NX, NY, NZ], [LX, LY, LZ])
    rock = og.make_rock(g, PERM, PORO)
    pv_r = rock.poro * g.cell_volumes
    N, T, _, _ = og.internal_connections(g, og.compute_trans(g, rock))
    ops = Ops(N, T, g.num_cells)
    nc = g.num_cells
    wc = 1 + NX * np.arange(1, 9) + NX * NY * 4
    w = og.add_well(None, g, rock, wc, "bhp", BHP_TARGET, radius=0.1, well_dir="y",
                    gravity=np.array([0.0, 0.0, og.NORM_GRAVITY]))[0]
    wi, dz = w.WI, w.dZ
    gradz = ops.grad(g.cell_centroids[:, 2])
    scatter = sp.csr_array((np.ones(wc.size), (wc, np.arange(wc.size))), shape=(nc, wc.size))

    rho = lambda p: RHO_REF * ad_exp(C_FLUID * (p - P_REF))
    pv = lambda p: pv_r * ad_exp(C_ROCK * (p - P_REF))

    def equations(p, bhp, qs, p0, dt):
        r = rho(p)
        flux = -(ops.T / MU) * (ops.grad(p) - gravity * ops.avg(r) * gradz)
        acc0 = pv_r * np.exp(C_ROCK * (p0 - P_REF)) * RHO_REF * np.exp(C_FLUID * (p0 - P_REF))
        mass = (1.0 / dt) * (pv(p) * r - acc0) + ops.div(ops.avg(r) * flux)
        # connection pressure bhp + g dz rho(bhp); inflow rate WI rho(p)/mu (p_conn - p)
        p_conn = bhp[np.zeros(wc.size, dtype=np.int64)] + gravity * dz * rho(bhp)[np.zeros(wc.size, dtype=np.int64)]
        q = wi * (rho(p[wc]) / MU) * (p_conn - p[wc])
        mass = mass - ad_matmul(scatter, q)
        return mass, qs - ad_sum(q) / RHO_SURFACE, bhp - BHP_TARGET

    p_init = hydrostatic_pressure(g.cell_centroids[:, 2], gravity)
    p, bhp, qs = p_init.copy(), np.array([p_init[wc[0]]]), np.array([0.0])
    dt = TOT_TIME / NUM_STEPS
    p_hist, bhp_hist, qs_hist = [p.copy()], [bhp[0]], [0.0]
    nits = []
    for step in range(steps):
        p0 = p.copy()
        done = False
        for nit in range(MAX_ITS):
            pa, ba, qa = init_variables_adi(p, bhp, qs)
            res, J = combine_equations(*equations(pa, ba, qa, p0, dt))
            upd = spla.spsolve(J.tocsc(), -res)
            p = p + upd[:nc]
            bhp = bhp + upd[nc:nc + 1]
            qs = qs + upd[nc + 1:]
            if np.linalg.norm(res) <= NEWTON_TOL:
                done = True
                break
        assert done, f"step {step + 1}: Newton did not converge"
        nits.append(nit + 1)
        p_hist.append(p.copy())
        bhp_hist.append(bhp[0])
        qs_hist.append(qs[0])
    return dict(p_hist=np.array(p_hist), bhp_hist=np.array(bhp_hist), qs_hist=np.array(qs_hist),
                p_init=p_init, WI=wi, dZ=dz, nits=np.array(nits), pv_r=pv_r)
