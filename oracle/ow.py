"""Fully-implicit oil-water model oracle (test infrastructure only).

Restates TwoPhaseOilWaterModelAD,
pymrst/modules/ad_blackoil/models/two_phase_oil_water_model.py:
  initial_state :172-186, _segment_mix_fractions :188-206, _topology_cdp
  :208-224, _compute_cdp_per_well :226-268, _well_contributions_batched
  :355-491, assemble :493-546, _converged :549-567, step :605-657,
and simulate_schedule_ad (pymrst/modules/ad_core/simulators/
simulate_schedule_ad.py:17-80), with the fluid of init_simple_adi_fluid
(pymrst/modules/ad_props/init_simple_adi_fluid.py:220-239,313-320) and the
TPFA operators of setup_operators_tpfa.py:158-209 / face_upstr.py:15-34.

One documented extension: ``c_r`` (rock compressibility) multiplies the pore
volume by pvMultR(p) = 1 + c_r (p - p_ref), as the AD-OO oil-water equations
do (pymrst/modules/ad_blackoil/utils/equations_oil_water.py:38-56,
black_oil_pore_volume.py:17-29).  ``c_r=None`` is the legacy model exactly.
"""
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .adi import ADI, ad_exp, ad_matmul, combine_equations, init_variables_adi, value
from .grid import BARSA, CartGrid, Rock, Well


@dataclass
class OWFluid:
    """init_simple_adi_fluid for phases 'WO' (init_simple_adi_fluid.py:94-322)."""
    mu: np.ndarray          # [mu_w, mu_o]
    rhoS: np.ndarray        # [rho_w, rho_o]
    n: np.ndarray           # Corey exponents
    c: np.ndarray           # compressibilities
    p_ref: float
    c_r: Optional[float] = None

    def b(self, i, p):
        if self.c[i] == 0.0:
            return p * 0.0 + 1.0
        return ad_exp(self.c[i] * (p - self.p_ref))

    def kr(self, i, s):
        return s ** self.n[i]

    def pv_mult(self, p):
        if self.c_r is None:
            return None
        return 1.0 + self.c_r * (p - self.p_ref)


@dataclass
class OWState:
    pressure: np.ndarray
    sW: np.ndarray
    qWs: np.ndarray
    qOs: np.ndarray
    bhp: np.ndarray

    def copy(self):
        return OWState(*(np.array(v, dtype=float, copy=True) for v in
                         (self.pressure, self.sW, self.qWs, self.qOs, self.bhp)))


class Ops:
    """TPFA operators on the internal connections N (setup_operators_tpfa.py:158-209)."""

    def __init__(self, N, T, nc):
        self.N, self.T, self.nc = N, T, nc
        nf = N.shape[0]
        self.C = sp.csr_array((np.tile([1.0, -1.0], nf), (np.repeat(np.arange(nf), 2), N.ravel())),
                              shape=(nf, nc))
        self.Ct = self.C.T.tocsr()
        self.M = sp.csr_array((np.full(2 * nf, 0.5), (np.repeat(np.arange(nf), 2), N.ravel())),
                              shape=(nf, nc))

    def grad(self, x):
        return -ad_matmul(self.C, x)

    def div(self, x):
        return ad_matmul(self.Ct, x)

    def avg(self, x):
        return ad_matmul(self.M, x)

    def upw(self, flag, x):
        up = self.N[:, 1].copy()
        up[flag] = self.N[flag, 0]
        if isinstance(x, ADI):
            return x[up]
        return np.asarray(x)[up]


def segment_mix_fractions(compi, cq_w, cq_o):
    """_segment_mix_fractions (two_phase_oil_water_model.py:188-206): the
    surface composition of the flow past each segment of a top-to-bottom
    perforation chain, i.e. of everything perforated at or below it (a
    reverse cumulative sum); segments with zero net flow take compi."""
    seg_w = np.cumsum(np.asarray(cq_w, float)[::-1])[::-1]
    seg_o = np.cumsum(np.asarray(cq_o, float)[::-1])[::-1]
    seg_t = seg_w + seg_o
    dead = seg_t == 0.0
    safe = np.where(dead, 1.0, seg_t)
    return np.where(dead, compi[0], seg_w / safe), np.where(dead, compi[1], seg_o / safe)


def topology_cdp(dZ, mix_w, mix_o, b_w_c, b_o_c, rhoS, gravity):
    """_topology_cdp (two_phase_oil_water_model.py:208-224): per-perforation
    mixture density sum(mix rhoS) / sum(mix / b), and the connection pressure
    drop cumsum(g rho_mix dz) down the chain from the reference depth."""
    vr = mix_w / b_w_c + mix_o / b_o_c
    rho_mix = (mix_w * rhoS[0] + mix_o * rhoS[1]) / vr
    dz = np.diff(np.concatenate(([0.0], np.asarray(dZ, float))))
    return np.cumsum(gravity * rho_mix * dz)


class OWModel:
    """TwoPhaseOilWaterModelAD restated (two_phase_oil_water_model.py:79-657)."""

    def __init__(self, g: CartGrid, rock: Rock, fluid: OWFluid, wells: Sequence[Well],
                 N, T, gravity: float = 0.0, ds_max=0.2, dp_max_rel=0.2,
                 tol_cnv=1e-3, tol_mb=1e-7, tol_well=1e-6, max_iterations=25,
                 linear_solver=None):
        self.g, self.rock, self.fluid = g, rock, fluid
        self.wells = list(wells)
        self.nc = g.num_cells
        self.ops = Ops(np.asarray(N), np.asarray(T, float), self.nc)
        self.pv = rock.poro * g.cell_volumes
        self.gravity = float(gravity)
        if gravity:
            z = g.cell_centroids[:, 2]
            self.gdz = gravity * (z[self.ops.N[:, 1]] - z[self.ops.N[:, 0]])
        else:
            self.gdz = np.zeros(self.ops.T.size)
        self.ds_max, self.dp_max_rel = ds_max, dp_max_rel
        self.tol_cnv, self.tol_mb, self.tol_well = tol_cnv, tol_mb, tol_well
        self.max_iterations = max_iterations
        self.linear_solver = linear_solver
        self._last_cdp = None
        self.last_linear_iterations = []

    n_phases = 2

    def initial_state(self, p0, sw0) -> OWState:
        """:172-186."""
        bhp = np.array([w.val if w.type == "bhp" else p0 for w in self.wells], float)
        qws = np.array([w.val * w.compi[0] if w.type == "rate" else 0.0 for w in self.wells], float)
        qos = np.array([w.val * w.compi[1] if w.type == "rate" else 0.0 for w in self.wells], float)
        return OWState(np.full(self.nc, float(p0)), np.full(self.nc, float(sw0)), qws, qos, bhp)

    # -- wells -------------------------------------------------------------
    def _perf_tables(self):
        wells = self.wells
        counts = np.array([w.cells.size for w in wells])
        return (counts, np.concatenate([w.cells for w in wells]).astype(np.int64),
                np.concatenate([w.WI for w in wells]), np.repeat(np.arange(len(wells)), counts))

    def well_contributions(self, p, mob_w, mob_o, b_w, b_o, qws, qos, bhp, cdp_per_well):
        """_well_contributions_batched (:355-491)."""
        wells = self.wells
        nw, nc = len(wells), self.nc
        counts, wc, tw, wop = self._perf_tables()
        n_perf = wc.size
        cdp = np.concatenate([np.asarray(c, float) for c in cdp_per_well])
        drawdown = -(bhp[wop] + cdp) + p[wc]
        inj = value(drawdown) < 0
        prod = ~inj
        ce_p = prod.astype(float) * tw
        cq_pw = -ce_p * mob_w[wc] * drawdown
        cq_po = -ce_p * mob_o[wc] * drawdown
        cq_psw = b_w[wc] * cq_pw
        cq_pso = b_o[wc] * cq_po
        grouping = sp.csr_array((np.ones(n_perf), (wop, np.arange(n_perf))), shape=(nw, n_perf))
        any_prod = np.zeros(nw, dtype=bool)
        np.logical_or.at(any_prod, wop, prod)
        wbq_w = qws * (value(qws) > 0).astype(float) - ad_matmul(grouping, cq_psw)
        wbq_o = qos * (value(qos) > 0).astype(float) - ad_matmul(grouping, cq_pso)
        wbqt = wbq_w + wbq_o
        dead = (value(wbqt) == 0.0) | (~any_prod)
        df = dead.astype(float)
        wbqt_safe = wbqt + df
        cw = np.array([w.compi[0] for w in wells])
        co = np.array([w.compi[1] for w in wells])
        mix_w = (wbq_w / wbqt_safe) * (1.0 - df) + cw * df
        mix_o = (wbq_o / wbqt_safe) * (1.0 - df) + co * df
        mt = mob_w[wc] + mob_o[wc]
        ce_i = inj.astype(float) * tw
        cqt_i = -ce_i * (mt * drawdown)
        vol_rat = mix_w[wop] / b_w[wc] + mix_o[wop] / b_o[wc]
        cqt_is = cqt_i / vol_rat
        cq_sw = cq_psw + mix_w[wop] * cqt_is
        cq_so = cq_pso + mix_o[wop] * cqt_is
        scatter = sp.csr_array((np.ones(n_perf), (wc, np.arange(n_perf))), shape=(nc, n_perf))
        water_adj = ad_matmul(scatter, cq_sw)
        oil_adj = ad_matmul(scatter, cq_so)
        eq_w = qws - ad_matmul(grouping, cq_sw)
        eq_o = qos - ad_matmul(grouping, cq_so)
        bm = np.array([w.type == "bhp" for w in wells], float)
        wv = np.array([w.val for w in wells])
        ctrl = (bhp - wv) * bm + (qws + qos - wv) * (1.0 - bm)
        return water_adj, oil_adj, eq_w, eq_o, ctrl, cq_sw, cq_so

    def compute_cdp_per_well(self, st0: OWState):
        """_compute_cdp_per_well (:226-268) with _segment_mix_fractions and
        _topology_cdp (:188-224)."""
        if not self.gravity:
            return [np.zeros(w.dZ.size) for w in self.wells]
        f = self.fluid
        p0 = st0.pressure
        b_w0, b_o0 = value(f.b(0, p0)), value(f.b(1, p0))
        mob_w0 = value(f.kr(0, st0.sW)) / f.mu[0]
        mob_o0 = value(f.kr(1, 1.0 - st0.sW)) / f.mu[1]
        prev = self._last_cdp or [np.zeros(w.dZ.size) for w in self.wells]
        *_, cq_sw, cq_so = self.well_contributions(
            p0, mob_w0, mob_o0, b_w0, b_o0, st0.qWs, st0.qOs, st0.bhp, prev)
        cq_sw, cq_so = value(cq_sw), value(cq_so)
        out, off = [], 0
        for w in self.wells:
            n = w.cells.size
            mix_w, mix_o = segment_mix_fractions(w.compi, cq_sw[off:off + n], cq_so[off:off + n])
            off += n
            out.append(topology_cdp(w.dZ, mix_w, mix_o, b_w0[w.cells], b_o0[w.cells], f.rhoS, self.gravity))
        self._last_cdp = out
        return out

    # -- equations -----------------------------------------------------------
    def assemble(self, st0: OWState, p, sw, qws, qos, bhp, dt, cdp_per_well=None):
        """assemble (:493-546)."""
        f, ops, pv = self.fluid, self.ops, self.pv
        b_w, b_o = f.b(0, p), f.b(1, p)
        so = 1.0 - sw
        mob_w = f.kr(0, sw) / f.mu[0]
        mob_o = f.kr(1, so) / f.mu[1]
        b_w0, b_o0 = f.b(0, st0.pressure), f.b(1, st0.pressure)
        dp = ops.grad(p)
        rho_w, rho_o = f.rhoS[0] * b_w, f.rhoS[1] * b_o
        pot_w = dp - ops.avg(rho_w) * self.gdz
        pot_o = dp - ops.avg(rho_o) * self.gdz
        flag_w = value(pot_w) <= 0
        flag_o = value(pot_o) <= 0
        v_w = -ops.T * ops.upw(flag_w, mob_w) * pot_w
        v_o = -ops.T * ops.upw(flag_o, mob_o) * pot_o
        pvm, pvm0 = f.pv_mult(p), f.pv_mult(st0.pressure)
        if pvm is None:
            water = (1.0 / dt) * (pv * b_w * sw - pv * b_w0 * st0.sW)
            oil = (1.0 / dt) * (pv * b_o * so - pv * b_o0 * (1.0 - st0.sW))
        else:
            water = (1.0 / dt) * (pv * pvm * b_w * sw - pv * pvm0 * b_w0 * st0.sW)
            oil = (1.0 / dt) * (pv * pvm * b_o * so - pv * pvm0 * b_o0 * (1.0 - st0.sW))
        water = water + ops.div(ops.upw(flag_w, b_w) * v_w)
        oil = oil + ops.div(ops.upw(flag_o, b_o) * v_o)
        if cdp_per_well is None:
            cdp_per_well = [np.zeros(w.dZ.size) for w in self.wells]
        wa, oa, eq_w, eq_o, ctrl, _, _ = self.well_contributions(
            p, mob_w, mob_o, b_w, b_o, qws, qos, bhp, cdp_per_well)
        return water - wa, oil - oa, eq_w, eq_o, ctrl

    def residual_and_jacobian(self, st0, st, dt, cdp_per_well=None):
        """(r, J) in the reference's variable-blocked ordering
        [water(nc), oil(nc), eq_w(nw), eq_o(nw), ctrl(nw)] x [p, sW, qWs, qOs, bhp]."""
        v = init_variables_adi(st.pressure, st.sW, st.qWs, st.qOs, st.bhp)
        eqs = self.assemble(st0, *v, dt, cdp_per_well)
        r, J = combine_equations(*eqs)
        return r, J, eqs

    def converged(self, eqs, dt, st: OWState):
        """_converged (:549-567) with cnv_mb_convergence (convergence.py:44-69)."""
        water, oil, eq_w, eq_o, ctrl = (value(e) for e in eqs)
        f, pv = self.fluid, self.pv
        bw = np.mean(1.0 / value(f.b(0, st.pressure)))
        bo = np.mean(1.0 / value(f.b(1, st.pressure)))
        cnv = max(bw * dt * np.abs(water / pv).max(), bo * dt * np.abs(oil / pv).max())
        mb = max(bw * dt * abs(water.sum()) / pv.sum(), bo * dt * abs(oil.sum()) / pv.sum())
        rate = [abs(w.val) for w in self.wells if w.type == "rate"]
        q_ref = max(max(rate), 1e-12) if rate else 1.0
        well_ok = (np.abs(eq_w).max() <= self.tol_well * q_ref
                   and np.abs(eq_o).max() <= self.tol_well * q_ref)
        ctrl_ok = np.all(np.abs(ctrl) <= self.tol_well *
                         np.array([BARSA if w.type == "bhp" else q_ref for w in self.wells]))
        return bool(cnv <= self.tol_cnv and mb <= self.tol_mb and well_ok and ctrl_ok), (cnv, mb)

    def solve_increment(self, r, J, st):
        """_solve_increment (:569-603)."""
        if self.linear_solver is None:
            return spla.spsolve(J.tocsc(), -r)
        x = self.linear_solver.solve(J.tocsr(), -r, self, {"pressure": st.pressure})
        self.last_linear_iterations.append(getattr(self.linear_solver, "last_iterations", 0))
        return x

    def apply_update(self, st: OWState, upd):
        """The update limits of step (:629-651)."""
        nc, nw = self.nc, len(self.wells)
        dp = upd[:nc]
        ds = upd[nc:2 * nc]
        cap = self.dp_max_rel * np.abs(st.pressure)
        dp = np.clip(dp, -cap, cap)
        ds = np.clip(ds, -self.ds_max, self.ds_max)
        st.pressure = st.pressure + dp
        st.sW = np.clip(st.sW + ds, 0.0, 1.0)
        st.qWs = st.qWs + upd[2 * nc:2 * nc + nw]
        st.qOs = st.qOs + upd[2 * nc + nw:2 * nc + 2 * nw]
        st.bhp = st.bhp + upd[2 * nc + 2 * nw:]

    def step(self, state: OWState, dt: float):
        """step (:605-657): returns (state, converged, iterations)."""
        st0 = state
        st = state.copy()
        cdp = self.compute_cdp_per_well(st0)
        for it in range(self.max_iterations):
            r, J, eqs = self.residual_and_jacobian(st0, st, dt, cdp)
            ok, _ = self.converged(eqs, dt, st)
            if ok:
                return st, True, it
            upd = self.solve_increment(r, J, st)
            self.apply_update(st, upd)
        r, J, eqs = self.residual_and_jacobian(st0, st, dt, cdp)
        return st, self.converged(eqs, dt, st)[0], self.max_iterations


def simulate_schedule(state0, model, timesteps, max_cuts=6):
    """simulate_schedule_ad.py:17-80: report steps with halving cuts."""
    state = state0
    states, its_per_step = [], np.zeros(len(timesteps), dtype=np.int64)
    for i, dt_report in enumerate(timesteps):
        t_local, its = 0.0, 0
        dt, cuts = float(dt_report), 0
        while t_local < dt_report - 1e-12 * dt_report:
            dt = min(dt, dt_report - t_local)
            cand, ok, nit = model.step(state, dt)
            its += nit
            if ok:
                state = cand
                t_local += dt
                dt = min(2.0 * dt, dt_report - t_local) if cuts else dt
            else:
                cuts += 1
                assert cuts <= max_cuts, f"step {i}: Newton failed after {max_cuts} cuts"
                dt /= 2.0
        its_per_step[i] = its
        states.append(state)
    return states, its_per_step
