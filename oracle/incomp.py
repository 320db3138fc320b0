"""incompTPFA oracle restatement (test infrastructure only).

Follows incomp_tpfa_ad, pymrst/modules/incomp/incomp_tpfa.py:252-389, with
its right-hand side _pressure_rhs (:392-435) and the well pressure drop of
pymrst/params/wells_and_bc/incomp_well_pressure_drop.py:9-26.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .grid import CartGrid, Well


@dataclass
class BC:
    face: np.ndarray
    type: List[str]
    value: np.ndarray


def add_bc(bc: Optional[BC], faces, kind, value) -> BC:
    faces = np.asarray(faces, dtype=np.int64).ravel()
    vals = np.broadcast_to(np.asarray(value, float), faces.shape).copy()
    if bc is None:
        return BC(faces, [kind] * faces.size, vals)
    return BC(np.concatenate([bc.face, faces]), bc.type + [kind] * faces.size,
              np.concatenate([bc.value, vals]))


@dataclass
class Src:
    cell: np.ndarray
    rate: np.ndarray


def add_source(src: Optional[Src], cells, rate) -> Src:
    cells = np.asarray(cells, dtype=np.int64).ravel()
    rate = np.broadcast_to(np.asarray(rate, float), cells.shape).copy()
    if src is None:
        return Src(cells, rate)
    return Src(np.concatenate([src.cell, cells]), np.concatenate([src.rate, rate]))


@dataclass
class IncompResult:
    pressure: np.ndarray
    flux: np.ndarray
    facePressure: np.ndarray
    well_flux: list
    well_pressure: list
    well_cdp: list
    A: object = None
    rhs: object = None


def incomp_well_pressure_drop(w: Well, mob: np.ndarray, rho: np.ndarray, norm_g: float):
    """pymrst/params/wells_and_bc/incomp_well_pressure_drop.py:9-26."""
    nph = rho.size
    injector = w.sign > 0 or (w.type == "rate" and w.val > 0)
    if injector:
        n = min(nph, w.compi.size)
        rho_mix = float(rho[:n] @ w.compi[:n])
    else:
        mw = mob[w.cells]
        frac = mw / mw.sum(axis=1, keepdims=True)
        n = min(nph, mob.shape[1])
        rho_mix = float((w.WI * (frac[:, :n] @ rho[:n])).sum() / w.WI.sum())
    return norm_g * w.dZ * rho_mix


def incomp_tpfa(g: CartGrid, t_half: np.ndarray, mob: np.ndarray, rho: np.ndarray,
                bc: Optional[BC] = None, src: Optional[Src] = None,
                wells: Optional[List[Well]] = None, gravity=None,
                solve=None) -> IncompResult:
    """incomp_tpfa_ad (incomp_tpfa.py:252-389).

    ``mob`` is (nc, nph) phase mobility, ``rho`` (nph,) densities.
    """
    wells = list(wells or [])
    gvec = np.zeros(3) if gravity is None else np.asarray(gravity, float).ravel()[:3]
    grav_on = np.linalg.norm(gvec) > 0
    nb = g.neighbors
    internal = np.all(nb >= 0, axis=1)
    cell_no = g.cell_no()
    cf = g.cell_faces
    nf, nc, nw = g.num_faces, g.num_cells, len(wells)
    n = nc + nw
    mob = np.atleast_2d(np.asarray(mob, float))
    totmob = mob.sum(axis=1)
    omega = (mob * rho).sum(axis=1) / totmob
    t_hf = t_half * totmob[cell_no]
    ft = 1.0 / np.bincount(cf, weights=1.0 / t_hf, minlength=nf)

    # _pressure_rhs (:392-435)
    if grav_on:
        cvec = g.face_centroids[cf] - g.cell_centroids[cell_no]
        grav_hf = omega[cell_no] * (cvec @ gvec)
    else:
        grav_hf = np.zeros(cf.size)
    ff = np.zeros(cf.size)
    gg = np.zeros(nc)
    hh = np.zeros(nf)
    d_face = np.zeros(nf, dtype=bool)
    d_val = np.zeros(nf)
    if src is not None and src.cell.size:
        gg += np.bincount(src.cell, weights=src.rate, minlength=nc)
    if bc is not None and bc.face.size:
        types = np.array(bc.type)
        ip = types == "pressure"
        d_face[bc.face[ip]] = True
        d_val[bc.face[ip]] = bc.value[ip]
        hd = d_face[cf]
        ff[hd] = -d_val[cf[hd]]
        ifl = types == "flux"
        hh[bc.face[ifl]] = -bc.value[ifl]

    sgn = np.where(nb[cf, 0] == cell_no, 1.0, -1.0)
    j = internal[cf] | d_face[cf]
    fg = np.bincount(cf[j], weights=grav_hf[j] * sgn[j], minlength=nf)
    rhs = np.zeros(n)
    rhs[:nc] += np.bincount(cell_no, weights=-ft[cf] * (sgn * fg[cf] + ff), minlength=nc)
    rhs[:nc] += gg
    rhs[:nc] += np.bincount(cell_no, weights=-hh[cf], minlength=nc)

    d = np.zeros(nc)
    c_rows, c_cols, c_vals = [], [], []
    d_well = np.zeros(nw)
    cdp = []
    for k, w in enumerate(wells):
        wc = w.cells
        wi = w.WI * totmob[wc]
        dp = incomp_well_pressure_drop(w, mob, rho, float(np.linalg.norm(gvec)))
        cdp.append(dp)
        d[wc] += wi
        if w.type == "bhp":
            ww = wi.max()
            rhs[nc + k] += ww * w.val
            rhs[wc] += wi * (w.val + dp)
            d_well[k] = ww
        elif w.type == "rate":
            rhs[nc + k] += w.val
            rhs[wc] += wi * dp
            c_rows += [k] * wc.size
            c_cols += wc.tolist()
            c_vals += (-wi).tolist()
            d_well[k] = wi.sum()
            rhs[nc + k] -= wi @ dp
        else:
            raise ValueError(w.type)
    hd = d_face[cf]
    d += np.bincount(cell_no[hd], weights=t_hf[hd], minlength=nc)
    d += np.bincount(nb[internal].ravel(order="F"), weights=np.tile(ft[internal], 2), minlength=nc)
    ni = nb[internal]
    rows = np.concatenate([ni[:, 0], ni[:, 1], np.arange(nc)])
    cols = np.concatenate([ni[:, 1], ni[:, 0], np.arange(nc)])
    vals = np.concatenate([-ft[internal], -ft[internal], d])
    a = sp.coo_array((vals, (rows, cols)), shape=(nc, nc)).tocsr()
    if nw:
        c = sp.coo_array((c_vals, (c_rows, c_cols)), shape=(nw, nc)).tocsr()
        a = sp.block_array([[a, c.T], [c, sp.diags_array(d_well, format="csr")]], format="csr")
    if not np.any(d_face) and not any(w.type == "bhp" for w in wells):
        a = a.tolil()
        if a[0, 0] > 0:
            a[0, 0] = 2.0 * a[0, 0]
        else:
            jm = int(np.argmax(a.diagonal()))
            a[jm, jm] = 2.0 * a[jm, jm]
        a = a.tocsr()
    p = spla.spsolve(a.tocsc(), rhs) if solve is None else solve(a, rhs)

    fpress = (np.bincount(cf, weights=(p[cell_no] + grav_hf) * t_hf, minlength=nf)
              / np.bincount(cf, weights=t_hf, minlength=nf))
    bnd = np.any(nb == -1, axis=1)
    fpress[bnd] -= hh[bnd] / ft[bnd]
    fpress[d_face] = d_val[d_face]
    flux = np.zeros(nf)
    flux[internal] = -ft[internal] * (p[ni[:, 1]] - p[ni[:, 0]] - fg[internal])
    bsgn = np.where(nb[~internal, 1] == -1, 1.0, -1.0)
    bcell = nb[~internal].sum(axis=1) + 1
    fg_all = np.bincount(cf, weights=grav_hf, minlength=nf)
    flux[~internal] = -bsgn * ft[~internal] * (fpress[~internal] - p[bcell] - fg_all[~internal])
    wflux = [w.WI * totmob[w.cells] * (p[nc + k] + cdp[k] - p[w.cells]) for k, w in enumerate(wells)]
    wp = [float(p[nc + k]) for k in range(nw)]
    return IncompResult(p[:nc].copy(), flux, fpress, wflux, wp, cdp, a, rhs)
