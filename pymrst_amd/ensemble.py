"""Batched oil-water ensembles: uneven realisations solved as one system.

The members' grids are joined into one disjoint union (cells and
connections offset, wells renumbered), so one assembly launch, one CPR setup
and one preconditioner or SpMV launch cover every member, and the AMG
hierarchy and line tables simply have one component per member.  FGMRES
keeps one Krylov recurrence per member (pmx_ow_set_segments): a single
Krylov space over the union stalls as members multiply, because one
polynomial has to damp every member's outlying eigenvalues.  Newton runs on the union: an
iteration is taken while any member is unconverged, and convergence is
judged member by member with the reference's CNV / MB / well tests
(two_phase_oil_water_model.py _converged :549-567), so a member is never
declared converged on the strength of the others.  A time-step cut halves
the step of the whole ensemble.
"""
from dataclasses import dataclass, replace
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .grid import BARSA, Rock, add_well, cart_grid, cartesian_lines, compute_trans, make_rock, tpfa_connections
from .ow import CPRSolver, TwoPhaseOilWaterModelAD, init_simple_adi_fluid


@dataclass
class UnionGrid:
    """What the oil-water model reads of a grid, for a union of grids."""
    num_cells: int
    cell_centroids: np.ndarray
    cell_volumes: np.ndarray
    cartDims: Optional[tuple] = None


class EnsembleModel(TwoPhaseOilWaterModelAD):
    """TwoPhaseOilWaterModelAD over the union of the members, with
    member-by-member convergence."""

    def setup_members(self, cell_member: np.ndarray, well_member: np.ndarray, nmem: int):
        dev = self.device
        self.nmem = nmem
        self.cell_member = torch.as_tensor(cell_member, dtype=torch.long, device=dev)
        self.well_member = torch.as_tensor(well_member, dtype=torch.long, device=dev)
        self.pv_t = torch.as_tensor(self.pv, dtype=torch.float64, device=dev)
        self.nc_member = torch.bincount(self.cell_member, minlength=nmem).to(torch.float64)
        self.pv_member = torch.zeros(nmem, dtype=torch.float64, device=dev).index_add_(0, self.cell_member,
                                                                                          self.pv_t)
        qref = np.zeros(nmem)
        has_rate = np.zeros(nmem, bool)
        for k, w in enumerate(self.wells):
            if w.type == "rate":
                m = well_member[k]
                qref[m] = max(qref[m], abs(w.val))
                has_rate[m] = True
        qref = np.where(has_rate, np.maximum(qref, 1e-12), 1.0)
        self.qref_member = torch.as_tensor(qref, device=dev)
        self.well_is_bhp = torch.as_tensor([w.type == "bhp" for w in self.wells], device=dev)
        # unknown -> member, for freezing converged members
        self.unknown_member = torch.cat([self.cell_member.repeat_interleave(2),
                                         self.well_member.repeat_interleave(3)])
        self.member_its = np.zeros(nmem, dtype=np.int64)
        self._ok_m = None

    def _seg(self, v, op, idx, n):
        out = torch.zeros(n, dtype=v.dtype, device=v.device)
        return out.scatter_reduce(0, idx, v, reduce=op, include_self=False)

    def converged(self, dt: float, st):
        f = self.fluid
        r = self.residual_device()
        nc, M = self.nc, self.nmem
        rw, ro = r[0:2 * nc:2], r[1:2 * nc:2]
        p = st.p
        inv_bw = torch.exp(-f.c[0] * (p - f.p_ref)) if f.c[0] else torch.ones_like(p)
        inv_bo = torch.exp(-f.c[1] * (p - f.p_ref)) if f.c[1] else torch.ones_like(p)
        cm = self.cell_member
        Bw = self._seg(inv_bw, "sum", cm, M) / self.nc_member
        Bo = self._seg(inv_bo, "sum", cm, M) / self.nc_member
        cnv = torch.maximum(Bw * dt * self._seg((rw / self.pv_t).abs(), "amax", cm, M),
                            Bo * dt * self._seg((ro / self.pv_t).abs(), "amax", cm, M))
        mb = torch.maximum(Bw * dt * self._seg(rw, "sum", cm, M).abs() / self.pv_member,
                           Bo * dt * self._seg(ro, "sum", cm, M).abs() / self.pv_member)
        ok_m = (cnv <= self.tol_cnv) & (mb <= self.tol_mb)
        if self.nw:
            wr = r[2 * nc:].view(-1, 3)
            q = self.qref_member[self.well_member]
            well_ok = (wr[:, 0].abs() <= self.tol_well * q) & (wr[:, 1].abs() <= self.tol_well * q)
            scale = torch.where(self.well_is_bhp, torch.full_like(q, BARSA), q)
            well_ok &= wr[:, 2].abs() <= self.tol_well * scale
            bad = self._seg((~well_ok).to(torch.float64), "amax", self.well_member, M)
            ok_m &= bad == 0
        self.unconverged_members = int((~ok_m).sum())
        self.last_cnv_mb = (float(cnv.max()), float(mb.max()))
        self._ok_m = ok_m
        return bool(ok_m.all()), float(p.mean())

    def solve_and_update(self, st, pmean: float) -> bool:
        """Members that have converged keep their state: their residual rows
        are zeroed, so FGMRES gives them a zero update (as if each member ran
        its own Newton loop), and only the others count an iteration."""
        if self._ok_m is not None:
            mask = (~self._ok_m)[self.unknown_member].to(torch.float64)
            _lib.check(self.lib.pmx_ow_mask_residual(self.h, _lib.ptr(mask)), "pmx_ow_mask_residual")
            self.member_its += (~self._ok_m).cpu().numpy()
        return super().solve_and_update(st, pmean)

    def reset(self):
        super().reset()
        if hasattr(self, "member_its"):
            self.member_its[:] = 0


def build_ensemble(cases: Sequence, linear_solver: Optional[CPRSolver] = None, device=None,
                   segmented: bool = True) -> tuple:
    """(EnsembleModel, initial host state, member cell ranges) for a list of
    OWCase members sharing one fluid (the first member's).  With
    `segmented`, FGMRES keeps one Krylov recurrence per member."""
    c0 = cases[0]
    fluid = init_simple_adi_fluid(c0.mu, c0.rho, c0.n, c0.c, c0.p_ref, c_r=c0.c_r)
    Ns, Ts, cents, vols, perms, poros, wells, lines_ptr, lines_cells = [], [], [], [], [], [], [], [0], []
    cell_member, well_member, ranges = [], [], []
    off = 0
    for m, c in enumerate(cases):
        g = cart_grid(c.dims, c.phys)
        rock = make_rock(g, c.perm, c.poro)
        N, T, _, _ = tpfa_connections(g, compute_trans(g, rock))
        Ns.append(N + off)
        Ts.append(T)
        cents.append(g.cell_centroids)
        vols.append(g.cell_volumes)
        perms.append(rock.perm)
        poros.append(rock.poro)
        grav = np.array([0.0, 0.0, c.gravity]) if c.gravity else None
        mw = None
        for cells, kind, val, compi in c.wells:
            mw = add_well(mw, g, rock, cells, kind, val, compi=np.array(compi), sign=1 if kind == "rate" else -1,
                          gravity=grav)
        for w in mw:
            wells.append(replace(w, cells=w.cells + off, name=f"M{m}_{w.name}"))
            well_member.append(m)
        axis, ptr, lc = cartesian_lines(g, N, T)
        lines_cells.append(lc + off)
        lines_ptr.extend((ptr[1:] + lines_ptr[-1]).tolist())
        cell_member.append(np.full(g.num_cells, m))
        ranges.append((off, off + g.num_cells))
        off += g.num_cells
    ug = UnionGrid(num_cells=off, cell_centroids=np.vstack(cents), cell_volumes=np.concatenate(vols))
    rock = Rock(perm=np.vstack(perms), poro=np.concatenate(poros))
    model = EnsembleModel(ug, rock, fluid, wells, gravity=c0.gravity, linear_solver=linear_solver, device=device,
                          N=np.vstack(Ns), T=np.concatenate(Ts))
    ptr = np.ascontiguousarray(lines_ptr, dtype=np.int32)
    cells = np.ascontiguousarray(np.concatenate(lines_cells), dtype=np.int32)
    _lib.check(model.lib.pmx_ow_set_lines(model.h, ptr.size - 1, _lib.ptr(ptr), _lib.ptr(cells)),
               "pmx_ow_set_lines")
    model.setup_members(np.concatenate(cell_member), np.asarray(well_member), len(cases))
    if segmented:
        cptr = np.ascontiguousarray([0] + [b for _, b in ranges], dtype=np.int32)
        wptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum(np.bincount(well_member, minlength=len(cases)))]),
                                    dtype=np.int32)
        _lib.check(model.lib.pmx_ow_set_segments(model.h, len(cases), _lib.ptr(cptr), _lib.ptr(wptr)),
                   "pmx_ow_set_segments")
    st = model.initial_state(c0.p0, c0.sw0)
    for c, (a, b) in zip(cases, ranges):
        if c.p_init is not None:
            st.pressure[a:b] = c.p_init
    return model, st, ranges
