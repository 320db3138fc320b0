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
declared converged on the strength of the others.

simulate_ensemble gives every member its own time stepping: each member
has its own ministep length (per-cell step lengths in the assembly,
pmx_ow_set_cell_dt), accepts its ministeps, is cut and regrows by the rule
of simulate_schedule_ad (simulate_schedule_ad.py:17-80) on its own, and
stops when it reaches the report time; the union iterates until every
member has.  simulate_schedule_ad on the ensemble model instead shares one
step between all members (a cut for one member is taken by all).
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
        cell_member = np.asarray(cell_member, dtype=np.int64)
        if cell_member.size and np.any(np.diff(cell_member) < 0):
            raise ValueError("members' cells must be laid out one member after the other")
        self.cell_member = torch.as_tensor(cell_member, dtype=torch.long, device=dev)
        self.well_member = torch.as_tensor(well_member, dtype=torch.long, device=dev)
        self.well_member_np = np.asarray(well_member, dtype=np.int64)
        self.pv_t = torch.as_tensor(self.pv, dtype=torch.float64, device=dev)
        counts = np.bincount(cell_member, minlength=nmem)
        self.member_cell_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.nc_member = counts.astype(np.float64)
        pv = np.asarray(self.pv, dtype=np.float64)
        self.pv_member = np.array([pv[a:b].sum() for a, b in zip(self.member_cell_ptr[:-1],
                                                                  self.member_cell_ptr[1:])])
        _lib.check(self.lib.pmx_ow_set_members(self.h, nmem, _lib.ptr(self.member_cell_ptr)), "pmx_ow_set_members")
        qref = np.zeros(nmem)
        has_rate = np.zeros(nmem, bool)
        for k, w in enumerate(self.wells):
            if w.type == "rate":
                m = well_member[k]
                qref[m] = max(qref[m], abs(w.val))
                has_rate[m] = True
        self.qref_member = np.where(has_rate, np.maximum(qref, 1e-12), 1.0)
        self.well_is_bhp = np.array([w.type == "bhp" for w in self.wells], dtype=bool)
        # unknown -> member, for freezing converged members
        self.unknown_member = torch.cat([self.cell_member.repeat_interleave(2),
                                         self.well_member.repeat_interleave(3)])
        self.member_its = np.zeros(nmem, dtype=np.int64)
        self._ok_m = None
        self._ok_np = None
        self._frozen = None
        perf_member = np.repeat(np.asarray(well_member), [w.cells.size for w in self.wells]) \
            if self.nw else np.zeros(0, np.int64)
        self.perf_member = torch.as_tensor(perf_member, dtype=torch.long, device=dev)

    member_dt = None      # per-member step lengths (torch, nmem) while simulate_ensemble runs

    def converged(self, dt: float, st):
        """The single model's criterion (_converged) for every member on its
        own cells, wells and step length, from one device reduction per
        member (pmx_ow_member_convergence); members already at the report
        time count as converged."""
        M = self.nmem
        dt = self._member_dt_np if self.member_dt is not None else np.full(M, float(dt))
        out = np.zeros((M, 7))
        wr = np.zeros(3 * self.nw)
        _lib.check(self.lib.pmx_ow_member_convergence(self.h, _lib.ptr(st.p), _lib.ptr(out),
                                                      _lib.ptr(wr) if self.nw else None),
                   "pmx_ow_member_convergence")
        Bw = out[:, 4] / self.nc_member
        Bo = out[:, 5] / self.nc_member
        cnv = np.maximum(Bw * dt * out[:, 0], Bo * dt * out[:, 1])
        mb = np.maximum(Bw * dt * np.abs(out[:, 2]) / self.pv_member, Bo * dt * np.abs(out[:, 3]) / self.pv_member)
        ok_m = (cnv <= self.tol_cnv) & (mb <= self.tol_mb)
        if self.nw:
            wr = wr.reshape(-1, 3)
            q = self.qref_member[self.well_member_np]
            well_ok = (np.abs(wr[:, 0]) <= self.tol_well * q) & (np.abs(wr[:, 1]) <= self.tol_well * q)
            scale = np.where(self.well_is_bhp, BARSA, q)
            well_ok &= np.abs(wr[:, 2]) <= self.tol_well * scale
            bad = np.zeros(M, dtype=bool)
            np.logical_or.at(bad, self.well_member_np, ~well_ok)
            ok_m &= ~bad
        if self._frozen is not None:      # members already at the report time stay put
            ok_m |= self._frozen
        self.unconverged_members = int((~ok_m).sum())
        self.last_cnv_mb = (float(cnv.max()), float(mb.max()))
        self._ok_np = ok_m
        self._ok_m = torch.as_tensor(ok_m, device=self.device)
        return bool(ok_m.all()), float(out[:, 6].sum() / self.nc)

    def solve_and_update(self, st, pmean: float) -> bool:
        """Members that have converged keep their state: their residual rows
        are zeroed, so FGMRES gives them a zero update (as if each member ran
        its own Newton loop), and only the others count an iteration."""
        if self._ok_m is not None:
            mask = (~self._ok_m)[self.unknown_member].to(torch.float64)
            _lib.check(self.lib.pmx_ow_mask_residual(self.h, _lib.ptr(mask)), "pmx_ow_mask_residual")
            self.member_its += ~self._ok_np
        return super().solve_and_update(st, pmean)

    def reset(self):
        super().reset()
        if hasattr(self, "member_its"):
            self.member_its[:] = 0
        self.member_dt = None
        self._frozen = None
        _lib.check(self.lib.pmx_ow_set_cell_dt(self.h, _lib.ptr(None)), "pmx_ow_set_cell_dt")

    # -- per-member time stepping ----------------------------------------------
    def _set_member_dt(self, dtm: np.ndarray):
        self._member_dt_np = np.array(dtm, dtype=np.float64)
        self.member_dt = torch.as_tensor(dtm, dtype=torch.float64, device=self.device)
        self._dt_cell = self.member_dt[self.cell_member].contiguous()   # kept alive: the library reads it
        _lib.check(self.lib.pmx_ow_set_cell_dt(self.h, _lib.ptr(self._dt_cell)), "pmx_ow_set_cell_dt")

    def _copy_members(self, dst, src, sel: np.ndarray):
        """dst[member slices] = src[member slices] for the members in sel."""
        m = torch.as_tensor(sel, device=self.device)
        cm, wm = m[self.cell_member], m[self.well_member]
        dst.p[cm] = src.p[cm]
        dst.sw[cm] = src.sw[cm]
        if self.nw:
            dst.qws[wm] = src.qws[wm]
            dst.qos[wm] = src.qos[wm]
            dst.bhp[wm] = src.bhp[wm]

    def _update_cdp(self, st0, sel: np.ndarray):
        """Well heads from state0 for the members in sel (once per ministep,
        as the member alone computes them); the others keep theirs."""
        if not self.gravity or not self.nw:
            return None
        old = self._cdp_dev.clone()
        self.compute_cdp_per_well(st0)
        keep = ~torch.as_tensor(sel, device=self.device)[self.perf_member]
        n = keep.numel()
        self._cdp_dev[:n][keep] = old[:n][keep]
        return self._cdp_dev


def _member_ranges(model):
    """Per member: cell, well and perforation ranges in the union (members
    are laid out one after the other)."""
    if getattr(model, "_ranges_cache", None) is None:
        cm = model.cell_member.cpu().numpy()
        wm = model.well_member.cpu().numpy()
        pm = model.perf_member.cpu().numpy()
        def ptr(a):
            return np.concatenate([[0], np.cumsum(np.bincount(a, minlength=model.nmem))]).astype(np.int64)
        model._ranges_cache = (ptr(cm), ptr(wm), ptr(pm))
    return model._ranges_cache


def _transfer(src, s_states, dst, d_states, members, back=False):
    """Copy the member slices of states and of the well heads between a
    union and a union of a subset: members[j] of src is member j of dst.
    back copies dst -> src."""
    sc, sw_, sp = _member_ranges(src)
    dc, dw, dp = _member_ranges(dst)
    for j, k in enumerate(members):
        for a, b in zip(s_states, d_states):
            x, y = (b, a) if back else (a, b)
            xs = (dc[j], dc[j + 1]) if back else (sc[k], sc[k + 1])
            ys = (sc[k], sc[k + 1]) if back else (dc[j], dc[j + 1])
            y.p[ys[0]:ys[1]] = x.p[xs[0]:xs[1]]
            y.sw[ys[0]:ys[1]] = x.sw[xs[0]:xs[1]]
            xw = (dw[j], dw[j + 1]) if back else (sw_[k], sw_[k + 1])
            yw = (sw_[k], sw_[k + 1]) if back else (dw[j], dw[j + 1])
            for f in ("qws", "qos", "bhp"):
                getattr(y, f)[yw[0]:yw[1]] = getattr(x, f)[xw[0]:xw[1]]
        # the well heads travel too: the next ministep's heads start from them
        if src.gravity and src.nw:
            if back:
                src._cdp_dev[sp[k]:sp[k + 1]] = dst._cdp_dev[dp[j]:dp[j + 1]]
            else:
                dst._cdp_dev[dp[j]:dp[j + 1]] = src._cdp_dev[sp[k]:sp[k + 1]]
    if src.gravity and src.nw:
        (src if back else dst)._last_cdp = []     # the heads above are the members' current ones


def _run_level(model, st0, st, D, t, dtm, cuts, it, its_step, restart, vol, max_cuts, compact, observer, i):
    """Newton iterations of one union until every member reaches the report
    time D; per-member arrays are updated in place (vol, (M, 2): the surface
    water and oil each member's wells have moved, summed over its accepted
    ministeps as dt times the converged rates).  Returns the union iteration
    count (this level and the levels below)."""
    M = model.nmem
    done = np.zeros(M, dtype=bool)
    tol = 1e-12
    n_union = 0
    cells = np.diff(_member_ranges(model)[0])
    cdp = model._cdp_dev if (model.gravity and model.nw) else None
    model._set_member_dt(dtm)      # members may arrive in the middle of a ministep
    while not done.all():
        act = ~done
        if compact and cells[act].sum() < compact * cells.sum() and act.sum() < M:
            # the members still iterating move to a union of their own
            members = np.flatnonzero(act)
            sub, _, _ = build_ensemble([model.cases[k] for k in members], model.linear_solver, model.device,
                                       segmented=model.segmented, pieces=[model.pieces[k] for k in members])
            for attr in ("max_iterations", "tol_cnv", "tol_mb", "tol_well", "ds_max", "dp_max_rel"):
                setattr(sub, attr, getattr(model, attr))
            s0s = sub.device_state(sub.initial_state(model.cases[0].p0, model.cases[0].sw0))
            s1s = s0s.copy()
            _transfer(model, (st0, st), sub, (s0s, s1s), members)
            sl = [a[members].copy() for a in (t, dtm, cuts, it, its_step, restart, vol)]
            n_union += _run_level(sub, s0s, s1s, D, *sl, max_cuts, compact, observer, i)
            _transfer(model, (st,), sub, (s1s,), members, back=True)
            for a, b in zip((t, dtm, cuts, it, its_step, vol), sl[:5] + sl[6:]):
                a[members] = b
            model.stats["newton"] += sub.stats["newton"]
            del sub
            break
        if restart.any():
            cdp = model._update_cdp(st0, restart)
            model._set_member_dt(dtm)
            restart[:] = False
        model._frozen = done.copy()
        model.assemble_device(st0, st, float(D), cdp)
        _, pmean = model.converged(float(D), st)
        ok = model._ok_np
        accept = ok & ~done
        if accept.any():
            q = torch.zeros((M, 2), dtype=torch.float64, device=model.device)
            q[:, 0].index_add_(0, model.well_member, st.qws)
            q[:, 1].index_add_(0, model.well_member, st.qos)
            vol[accept] += dtm[accept, None] * q.cpu().numpy()[accept]
        for m in np.flatnonzero(accept):     # accept the ministep
            t[m] += dtm[m]
            its_step[m] += it[m]
            if t[m] >= D * (1.0 - tol):
                done[m] = True
                continue
            dtm[m] = min(2.0 * dtm[m], D - t[m]) if cuts[m] else min(dtm[m], D - t[m])
            it[m] = 0
            restart[m] = True
        failed = ~ok & ~done & (it >= model.max_iterations)
        for m in np.flatnonzero(failed):          # cut from the ministep's start
            cuts[m] += 1
            if cuts[m] > max_cuts:
                raise RuntimeError(f"step {i}: member {m}: Newton failed after {max_cuts} timestep cuts")
            its_step[m] += it[m]
            it[m] = 0
            dtm[m] /= 2.0
            restart[m] = True
            if observer:
                observer("timestep_cut", {"report_step": i, "member": int(m), "dt": float(dtm[m])})
        if restart.any():
            accepted = restart & ~failed
            if accepted.any():
                model._copy_members(st0, st, accepted)
            if failed.any():
                model._copy_members(st, st0, failed)
            continue
        if done.all():
            break
        # one Newton update of the members still iterating (the others are frozen)
        model.stats["newton"] += 1
        n_union += 1
        model.solve_and_update(st, pmean)
        it[~ok & ~done] += 1
    model.member_dt = None
    model._frozen = None
    _lib.check(model.lib.pmx_ow_set_cell_dt(model.h, _lib.ptr(None)), "pmx_ow_set_cell_dt")
    return n_union


def simulate_ensemble(model: "EnsembleModel", state0, timesteps: Sequence[float], max_cuts: int = 6,
                      observer=None, keep_device: bool = False, compact: float = 0.2, volumes=None):
    """Report steps with per-member time stepping.  Each Newton iteration is
    one union assembly and one (segmented) solve over the members still
    iterating; a member that converges accepts its ministep and starts the
    next one from there (or stops at the report time); a member that reaches
    max_iterations is cut to half its step from its ministep's start.  When
    the members still iterating hold fewer than `compact` of the union's
    cells, they continue in a union of their own (0 disables), so a few slow
    members do not keep every iteration at the full ensemble's cost.
    Returns (states per report step, union Newton iterations per report
    step, Newton iterations per report step and member).  A list given as
    `volumes` receives, per report step, an (M, 2) array of the surface
    water and oil each member's wells moved over it (sum over the member's
    ministeps of dt times its converged well rates; injection positive)."""
    M = model.nmem
    state = model.device_state(state0)
    states, union_its, member_its = [], [], []
    for i, D in enumerate(np.atleast_1d(np.asarray(timesteps, dtype=float))):
        if observer:
            observer("report_step_started", {"report_step": i, "dt": float(D)})
        st0, st = state.copy(), state.copy()
        t, dtm = np.zeros(M), np.full(M, D)
        cuts, it, its_step = (np.zeros(M, dtype=np.int64) for _ in range(3))
        restart = np.ones(M, dtype=bool)
        vol = np.zeros((M, 2))
        n_union = _run_level(model, st0, st, D, t, dtm, cuts, it, its_step, restart, vol, max_cuts, compact,
                             observer, i)
        if volumes is not None:
            volumes.append(vol)
        state = st
        union_its.append(n_union)
        member_its.append(its_step)
        if observer:
            observer("step_accepted", {"report_step": i, "dt": float(D), "iterations": int(n_union),
                                       "member_cuts": int(cuts.sum())})
        states.append(state.copy() if keep_device else state.to_host())
    return states, np.array(union_its), np.array(member_its)


@dataclass
class MemberPieces:
    """A member's static data in its own cell numbering: connections,
    transmissibilities, geometry, rock, wells and lines."""
    nc: int
    N: np.ndarray
    T: np.ndarray
    centroids: np.ndarray
    volumes: np.ndarray
    perm: np.ndarray
    poro: np.ndarray
    wells: list
    line_ptr: np.ndarray
    line_cells: np.ndarray


def member_pieces(c) -> MemberPieces:
    g = cart_grid(c.dims, c.phys)
    rock = make_rock(g, c.perm, c.poro)
    N, T, _, _ = tpfa_connections(g, compute_trans(g, rock))
    grav = np.array([0.0, 0.0, c.gravity]) if c.gravity else None
    mw = None
    for cells, kind, val, compi in c.wells:
        mw = add_well(mw, g, rock, cells, kind, val, compi=np.array(compi), sign=1 if kind == "rate" else -1,
                      gravity=grav)
    _, ptr, lc = cartesian_lines(g, N, T)
    return MemberPieces(g.num_cells, N, T, g.cell_centroids, g.cell_volumes, rock.perm, rock.poro, list(mw or []),
                        np.asarray(ptr), np.asarray(lc))


def build_ensemble(cases: Sequence, linear_solver: Optional[CPRSolver] = None, device=None,
                   segmented: bool = True, pieces: Optional[Sequence[MemberPieces]] = None) -> tuple:
    """(EnsembleModel, initial host state, member cell ranges) for a list of
    OWCase members sharing one fluid (the first member's).  With
    `segmented`, FGMRES keeps one Krylov recurrence per member.  `pieces`
    (one MemberPieces per case, e.g. a larger union's `model.pieces`) skips
    recomputing the members' static data."""
    c0 = cases[0]
    fluid = init_simple_adi_fluid(c0.mu, c0.rho, c0.n, c0.c, c0.p_ref, c_r=c0.c_r)
    if pieces is None:
        pieces = [member_pieces(c) for c in cases]
    member_data = list(pieces)
    if len(member_data) != len(cases):
        raise ValueError("one MemberPieces per case")
    Ns, Ts, cents, vols, perms, poros, wells, lines_ptr, lines_cells = [], [], [], [], [], [], [], [0], []
    cell_member, well_member, ranges = [], [], []
    off = 0
    for m, (c, pc) in enumerate(zip(cases, member_data)):
        Ns.append(pc.N + off)
        Ts.append(pc.T)
        cents.append(pc.centroids)
        vols.append(pc.volumes)
        perms.append(pc.perm)
        poros.append(pc.poro)
        for w in pc.wells:
            wells.append(replace(w, cells=w.cells + off, name=f"M{m}_{w.name}"))
            well_member.append(m)
        lines_cells.append(pc.line_cells + off)
        lines_ptr.extend((pc.line_ptr[1:] + lines_ptr[-1]).tolist())
        cell_member.append(np.full(pc.nc, m))
        ranges.append((off, off + pc.nc))
        off += pc.nc
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
    model.cases = list(cases)
    model.pieces = member_data
    model.segmented = segmented
    st = model.initial_state(c0.p0, c0.sw0)
    for c, (a, b) in zip(cases, ranges):
        if c.p_init is not None:
            st.pressure[a:b] = c.p_init
    return model, st, ranges
