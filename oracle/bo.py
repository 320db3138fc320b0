"""Three-phase live-oil black-oil model oracle (test infrastructure only).

Restates ThreePhaseBlackOilModelAD,
pymrst/modules/ad_blackoil/models/three_phase_black_oil_model.py:
  assemble :314-381, _well :383-436, update_state :439-470, _flash :472-509,
  _conv_scaling / _converged :512-545, step :547-571,
with the flux functions of ad_blackoil/utils/get_flux_and_props_{water,oil,
gas}_bo.py (face density averaged, kr and mu upwinded separately on the flag
dp <= 0 -> N0, rho_o = b_o (rs rho_gS + rho_oS)), and the deck fluid of
ad_props/props/init_deck_adi_fluid.py:63-194: analytic PVTW and ROCK, PVDG
and the relative-permeability tables through interp_table
(utils/interp_table.py:17-59, MATLAB interp1 weights, linear extrapolation)
with extend_tab (ad_props/utils/extend_tab.py, flat rows one unit beyond),
PVTO through pvt_interp 'parallel' (ad_props/utils/interp_pvt.py:41-60,
175-230, slope form), and the saturation-weighted oil relperm
(init_deck_adi_fluid.py:52-57).  Cells use the reference's variable-blocked
order [p, sW, x] with x = rs (status 1) or sG (status 3).
"""
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np
import scipy.sparse.linalg as spla

from .adi import ADI, _rowscale, combine_equations, init_variables_adi, value
from .grid import BARSA, CartGrid, Rock, Well
from .ow import Ops

ETOL = float(np.sqrt(np.finfo(float).eps))


# -- table interpolation --------------------------------------------------------
def _interp1(xt, ft, x):
    """interp_table.py:20-30: f0 (1-t) + f1 t, linear extrapolation."""
    idx = np.clip(np.searchsorted(xt, x, side="right") - 1, 0, xt.size - 2)
    x0, x1, f0, f1 = xt[idx], xt[idx + 1], ft[idx], ft[idx + 1]
    t = (x - x0) / (x1 - x0)
    return f0 * (1 - t) + f1 * t, (f1 - f0) / (x1 - x0)


def _interp1_slope(xt, ft, x):
    """interp_pvt.py:55-62: f0 + slope (x - x0)."""
    idx = np.clip(np.searchsorted(xt, x, side="right") - 1, 0, xt.size - 2)
    x0, x1, f0, f1 = xt[idx], xt[idx + 1], ft[idx], ft[idx + 1]
    slope = (f1 - f0) / (x1 - x0)
    return f0 + slope * (x - x0), slope


def _chain(f, d, x):
    if isinstance(x, ADI):
        return ADI(f, [_rowscale(j, d) for j in x.jac])
    return f


def interp_table(xt, ft, x):
    xt, ft = np.asarray(xt, float), np.asarray(ft, float)
    if np.any(np.diff(xt) < 0):
        xt, ft = xt[::-1], ft[::-1]
    f, d = _interp1(xt, ft, value(x))
    return _chain(f, d, x)


@dataclass
class PVTTable:
    key: np.ndarray
    branches: list          # (p, F) per key
    sat_x: np.ndarray
    sat_f: np.ndarray


def make_pvt_table(key, branches) -> PVTTable:
    return PVTTable(np.asarray(key, float), branches, np.array([b[0][0] for b in branches]),
                    np.array([b[1][0] for b in branches]))


def _get_bins(r, v):
    lims = v.copy()
    lims[0], lims[-1] = -np.inf, np.inf
    b = np.clip(np.searchsorted(lims, r, side="right") - 1, 0, v.size - 2)
    return b, (r - v[b]) / (v[b + 1] - v[b]), 1.0 / (v[b + 1] - v[b])


def pvt_values(tab: PVTTable, p, rs, sat):
    """interp_pvt.py:175-230, 'parallel' strategy: (f, df/dp, df/drs)."""
    p, rs = np.atleast_1d(np.asarray(p, float)), np.atleast_1d(np.asarray(rs, float))
    f, dfdp, dfdr = np.zeros_like(p), np.zeros_like(p), np.zeros_like(p)
    sat = np.asarray(sat, bool).copy()
    sat[p < tab.sat_x[0]] = False
    if sat.any():
        f[sat], dfdp[sat] = _interp1_slope(tab.sat_x, tab.sat_f, p[sat])
    us = ~sat
    if us.any():
        pu, ru = p[us], rs[us]
        b, w, dwdr = _get_bins(ru, tab.key)
        ps, dpsdr = _interp1_slope(tab.key, tab.sat_x, ru)
        dp = pu - ps
        fl, fr, dl, dr = (np.zeros_like(pu) for _ in range(4))
        for tn in range(tab.key.size):
            il, ir = b == tn - 1, b == tn
            if not (il.any() or ir.any()):
                continue
            px, fx = tab.branches[tn]
            dpx = px - px[0]
            if il.any():
                fl[il], dl[il] = _interp1_slope(dpx, fx, dp[il])
            if ir.any():
                fr[ir], dr[ir] = _interp1_slope(dpx, fx, dp[ir])
        fu = fl * w + fr * (1 - w)
        dfu = dl * w + dr * (1 - w)
        f[us], dfdp[us], dfdr[us] = fu, dfu, (fl - fr) * dwdr - dfu * dpsdr
    return f, dfdp, dfdr


def pvt_interp(tab, p, rs, sat):
    f, dfdp, dfdr = pvt_values(tab, value(p), value(rs), sat)
    jac = None
    if isinstance(p, ADI):
        jac = [_rowscale(j, dfdp) for j in p.jac]
    if isinstance(rs, ADI):
        jr = [_rowscale(j, dfdr) for j in rs.jac]
        jac = jr if jac is None else [a + b for a, b in zip(jac, jr)]
    return f if jac is None else ADI(f, jac)


def extend_tab(tab, pm=1.0):
    tab = np.asarray(tab, float)
    first, last = tab[0].copy(), tab[-1].copy()
    first[0] -= pm
    last[0] += pm
    return np.vstack([first, tab, last])


# -- fluid ---------------------------------------------------------------------------
class BOFluid:
    """init_deck_blackoil_fluid (init_deck_adi_fluid.py:63-194) for PVTW,
    ROCK, PVDG, PVTO, SWOF, SGOF and surface densities."""

    def __init__(self, deck: dict):
        self.deck = deck
        self.rhoS = np.asarray(deck["rhoS"], float)
        self.w = deck["pvtw"]
        self.r = deck["rock_cp"]
        pvdg = np.asarray(deck["pvdg"], float)
        self.p_g = np.concatenate([[0.0], pvdg[:, 0]])
        self.bg_tab = np.concatenate([[1.0 / pvdg[0, 1]], 1.0 / pvdg[:, 1]])
        self.mug_tab = np.concatenate([[pvdg[0, 2]], pvdg[:, 2]])
        keys = np.array([k for k, _ in deck["pvto"]], float)
        branches = []
        for _, rows in deck["pvto"]:
            rows = np.asarray(rows, float)
            if rows.shape[0] == 1:
                rows = np.vstack([rows, rows[0]])
                rows[1, 0] = np.nextafter(rows[0, 0], np.inf)
            branches.append(rows)
        self.bo_tab = make_pvt_table(keys, [(r[:, 0], 1.0 / r[:, 1]) for r in branches])
        self.muo_tab = make_pvt_table(keys, [(r[:, 0], r[:, 2]) for r in branches])
        self.rs_p = np.concatenate([[0.0], [np.asarray(r, float)[0, 0] for _, r in deck["pvto"]]])
        self.rs_k = np.concatenate([[0.0], keys])
        swof, sgof = extend_tab(deck["swof"]), extend_tab(deck["sgof"])
        self.swcon = float(np.asarray(deck["swof"])[0, 0])
        self.sw_tab, self.krw_tab = swof[:, 0], swof[:, 1]
        self.sow_tab, self.krow_tab = 1.0 - swof[::-1, 0], swof[::-1, 2]
        self.sg_tab, self.krg_tab = sgof[:, 0], sgof[:, 1]
        self.sog_tab, self.krog_tab = 1.0 - (sgof[::-1, 0] + self.swcon), sgof[::-1, 2]

    def b_w(self, p):
        x = self.w["c"] * (p - self.w["p_ref"])
        return (1.0 + x + 0.5 * x * x) / self.w["b_ref"]

    def mu_w(self, p):
        return self.w["mu_ref"] + 0.0 * value(p)

    def pv_mult(self, p):
        x = self.r["c"] * (p - self.r["p_ref"])
        return 1.0 + x + 0.5 * x * x

    def b_o(self, p, rs, sat):
        return pvt_interp(self.bo_tab, p, rs, sat)

    def mu_o(self, p, rs, sat):
        return pvt_interp(self.muo_tab, p, rs, sat)

    def rs_sat(self, p):
        return interp_table(self.rs_p, self.rs_k, p)

    def b_g(self, p):
        return interp_table(self.p_g, self.bg_tab, p)

    def mu_g(self, p):
        return interp_table(self.p_g, self.mug_tab, p)

    def kr_w(self, s):
        return interp_table(self.sw_tab, self.krw_tab, s)

    def kr_g(self, s):
        return interp_table(self.sg_tab, self.krg_tab, s)

    def kr_o(self, sw, so, sg):
        """relPermWOG default (init_deck_adi_fluid.py:52-57)."""
        swcon = np.minimum(self.swcon, value(sw) - 1e-5)
        d = sg + sw - swcon
        ww = (sw - swcon) / d
        return (1.0 - ww) * interp_table(self.sog_tab, self.krog_tab, so) + ww * interp_table(
            self.sow_tab, self.krow_tab, so)


def phase_flux(ops: Ops, p_phase, rho, kr, mu, T, gdz):
    """get_flux_and_props_*_bo: dp = grad(p) - avg(rho) gdz, flag dp <= 0,
    kr and mu upwinded separately, v = -(kr_f/mu_f) T dp."""
    dp = ops.grad(p_phase) - ops.avg(rho) * gdz
    up = np.asarray(value(dp)) <= 0
    v = -(ops.upw(up, kr) / ops.upw(up, mu)) * T * dp
    return v, up, dp


@dataclass
class BOState:
    pressure: np.ndarray
    s: np.ndarray            # (nc, 3) water, oil, gas
    rs: np.ndarray
    status: np.ndarray       # 1 undersaturated, 3 saturated
    qWs: np.ndarray
    qOs: np.ndarray
    qGs: np.ndarray
    bhp: np.ndarray

    def copy(self):
        return BOState(*(np.array(v, copy=True) for v in (self.pressure, self.s, self.rs, self.status, self.qWs,
                                                           self.qOs, self.qGs, self.bhp)))


class BOModel:
    def __init__(self, g: CartGrid, rock: Rock, fluid: BOFluid, wells: Sequence[Well], N, T, gravity=0.0,
                 ds_max=0.2, tol_cnv=1e-3, tol_mb=1e-7, tol_well=1e-6, max_iterations=25):
        self.g, self.fluid, self.wells = g, fluid, list(wells)
        self.nc = g.num_cells
        self.ops = Ops(np.asarray(N), np.asarray(T, float), self.nc)
        self.pv_ref = rock.poro * g.cell_volumes
        z = g.cell_centroids[:, 2]
        self.gdz = gravity * (z[self.ops.N[:, 1]] - z[self.ops.N[:, 0]])
        self.ds_max, self.tol_cnv, self.tol_mb, self.tol_well = ds_max, tol_cnv, tol_mb, tol_well
        self.max_iterations = max_iterations

    def initial_state(self, p0, sw0, sg0, rs0=None) -> BOState:
        """Gas-free cells start undersaturated at rs0 (default rs_sat(p0));
        cells with free gas start saturated."""
        nc, nw = self.nc, len(self.wells)
        p = np.full(nc, float(p0)) if np.size(p0) == 1 else np.asarray(p0, float).copy()
        sw = np.broadcast_to(np.asarray(sw0, float), (nc,)).copy()
        sg = np.broadcast_to(np.asarray(sg0, float), (nc,)).copy()
        rsat = value(self.fluid.rs_sat(p))
        rs = rsat.copy() if rs0 is None else np.minimum(np.broadcast_to(np.asarray(rs0, float), (nc,)), rsat)
        status = np.where(sg > 0, 3, 1).astype(np.int64)
        s = np.column_stack([sw, 1.0 - sw - sg, sg])
        # rate wells start at their target split by compi, with the bhp 1 bar
        # beyond the perforated cells' pressures (every connection flows the
        # intended way; a rate injector that starts crossflowing injects the
        # produced oil in the first linearisation)
        bhp = np.array([w.val if w.type == "bhp" else
                        (p[w.cells].max() + 1e5 if w.val > 0 else p[w.cells].min() - 1e5) for w in self.wells], float)
        q = np.array([[w.val * float(c) for c in np.asarray(w.compi, float)[:3]] if w.type == "rate" else [0.0] * 3
                      for w in self.wells]).reshape(nw, 3)
        return BOState(p, s, rs, status, q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy(), bhp)

    # -- assembly (three_phase_black_oil_model.py:314-381) --------------------------
    def assemble(self, st0: BOState, st: BOState, p, sw, x, qws, qos, qgs, bhp, dt):
        f, ops = self.fluid, self.ops
        st1, st3 = st.status == 1, st.status == 3
        rs = x * st1.astype(float) + f.rs_sat(p) * (~st1).astype(float)
        sg = x * st3.astype(float)
        so = 1.0 - sw - sg
        b_w, mu_w = f.b_w(p), f.mu_w(p)
        b_o, mu_o = f.b_o(p, rs, st3), f.mu_o(p, rs, st3)
        b_g, mu_g = f.b_g(p), f.mu_g(p)
        rho_w = b_w * f.rhoS[0]
        rho_o = b_o * (rs * f.rhoS[2] + f.rhoS[1])
        rho_g = b_g * f.rhoS[2]
        kr_w, kr_o, kr_g = f.kr_w(sw), f.kr_o(sw, so, sg), f.kr_g(sg)
        T = ops.T
        v_w, up_w, _ = phase_flux(ops, p, rho_w, kr_w, mu_w, T, self.gdz)
        v_o, up_o, _ = phase_flux(ops, p, rho_o, kr_o, mu_o, T, self.gdz)
        v_g, up_g, _ = phase_flux(ops, p, rho_g, kr_g, mu_g, T, self.gdz)
        bWvW = ops.upw(up_w, b_w) * v_w
        bOvO = ops.upw(up_o, b_o) * v_o
        bGvG = ops.upw(up_g, b_g) * v_g
        rsbOvO = ops.upw(up_o, rs) * bOvO
        pv = self.pv_ref * f.pv_mult(p)
        p0, s0, rs0 = st0.pressure, st0.s, st0.rs
        pv0 = self.pv_ref * value(f.pv_mult(p0))
        b_w0 = value(f.b_w(p0))
        b_o0 = value(f.b_o(p0, rs0, st0.status == 3))
        b_g0 = value(f.b_g(p0))
        water = (1.0 / dt) * (pv * b_w * sw - pv0 * b_w0 * s0[:, 0]) + ops.div(bWvW)
        oil = (1.0 / dt) * (pv * b_o * so - pv0 * b_o0 * s0[:, 1]) + ops.div(bOvO)
        gas = (1.0 / dt) * (pv * (b_g * sg + rs * b_o * so) - pv0 * (b_g0 * s0[:, 2] + rs0 * b_o0 * s0[:, 1])) \
            + ops.div(bGvG + rsbOvO)
        mob = (kr_w / mu_w, kr_o / mu_o, kr_g / mu_g)
        eq = [[], [], []]
        ctrl = []
        for k, w in enumerate(self.wells):
            kk = np.array([k])
            cq, weqs, c = self._well(k, p, mob, (b_w, b_o, b_g), rs, (qws[kk], qos[kk], qgs[kk]), bhp[kk])
            wc = w.cells
            for ph, tgt in enumerate((water, oil, gas)):
                sub = -cq[ph]
                # scatter-add the connection rates into the cell rows
                import scipy.sparse as sps
                S = sps.csr_array((np.ones(wc.size), (wc, np.arange(wc.size))), shape=(self.nc, wc.size))
                delta = ADI(S @ sub.val, [(S @ j).tocsr() for j in sub.jac]) if isinstance(sub, ADI) else S @ sub
                if ph == 0:
                    water = water + delta
                elif ph == 1:
                    oil = oil + delta
                else:
                    gas = gas + delta
            for ph in range(3):
                eq[ph].append(weqs[ph])
            ctrl.append(c)
        return (water, oil, gas, _vcat(eq[0], water), _vcat(eq[1], water), _vcat(eq[2], water),
                _vcat(ctrl, water))

    def _well(self, k, p, mob, b, rs, q_s, bhp_k):
        """_well (:383-436): cdp = 0; producer/injector split with the
        wellbore mixture and dissolved-gas accounting."""
        w = self.wells[k]
        wc, tw = w.cells, np.asarray(w.WI, float)
        compi = np.asarray(w.compi, float)[:3]
        dd = -bhp_k + p[wc]
        inj = value(dd) < 0
        prod = ~inj
        mob_c = [m[wc] for m in mob]
        b_c = [x[wc] for x in b]
        rs_c = rs[wc]
        any_inj, any_prod = bool(inj.any()), bool(prod.any())
        if any_prod:
            con = prod * tw
            cq_p = [-(con * m) * dd for m in mob_c]
            cq_ps = [b_c[0] * cq_p[0], b_c[1] * cq_p[1], b_c[2] * cq_p[2] + rs_c * (b_c[1] * cq_p[1])]
            wbq = [q_s[ph] * float(value(q_s[ph])[0] > 0) - _sum(cq_ps[ph]) for ph in range(3)]
            wbqt = wbq[0] + wbq[1] + wbq[2]
            mix = list(compi) if value(wbqt)[0] == 0.0 else [wq / wbqt for wq in wbq]
        else:
            cq_ps = None
            mix = list(compi)
        if any_inj:
            mt = mob_c[0] + mob_c[1] + mob_c[2]
            cqt_i = -(inj * tw) * (mt * dd)
            vol_rat = mix[0] / b_c[0] + mix[1] / b_c[1] + (mix[2] - rs_c * mix[1]) / b_c[2]
            cqt_is = cqt_i / vol_rat
            cq_s = [(cq_ps[ph] if any_prod else 0.0) + mix[ph] * cqt_is for ph in range(3)]
        else:
            cq_s = cq_ps
        weqs = [q_s[ph] - _sum(cq_s[ph]) for ph in range(3)]
        if w.type == "bhp":
            c = bhp_k - w.val
        else:
            c = q_s[0] + q_s[1] + q_s[2] - w.val
        return cq_s, weqs, c

    def residual_and_jacobian(self, st0, st, dt):
        x = np.where(st.status == 1, st.rs, st.s[:, 2])
        v = init_variables_adi(st.pressure, st.s[:, 0], x, st.qWs, st.qOs, st.qGs, st.bhp)
        eqs = self.assemble(st0, st, *v, dt)
        r, J = combine_equations(*eqs)
        return r, J, eqs

    # -- update and flash (:439-509) --------------------------------------------------
    def update_state(self, st: BOState, upd) -> BOState:
        nc, nw = self.nc, len(self.wells)
        prev, new = st, st.copy()
        st1, st3 = prev.status == 1, prev.status == 3
        dp, dsw, dx = upd[:nc], upd[nc:2 * nc], upd[2 * nc:3 * nc]
        new.pressure = prev.pressure + dp
        dsg = np.where(st3, dx, 0.0)
        new.rs = np.maximum(prev.rs + np.where(st1, dx, 0.0), 0.0)
        ds = np.clip(np.column_stack([dsw, -(dsg + dsw), dsg]), -self.ds_max, self.ds_max)
        new.s = prev.s + ds
        off = 3 * nc
        new.qWs = prev.qWs + upd[off:off + nw]
        new.qOs = prev.qOs + upd[off + nw:off + 2 * nw]
        new.qGs = prev.qGs + upd[off + 2 * nw:off + 3 * nw]
        new.bhp = prev.bhp + upd[off + 3 * nw:off + 4 * nw]
        self.flash(new, prev, st1)
        return new

    def flash(self, st: BOState, prev: BOState, st1_prev):
        f = self.fluid
        sw, so, sg = st.s[:, 0].copy(), st.s[:, 1].copy(), st.s[:, 2].copy()
        rs = st.rs.copy()
        wat_only = sw > 1.0 - ETOL
        rs_sat = value(f.rs_sat(st.pressure))
        rs_sat0 = value(f.rs_sat(prev.pressure))
        gas = ((sg > 0) | (rs == 0)) & ~st1_prev
        gas |= wat_only
        gas |= (sg < 0) & (prev.s[:, 2] > ETOL)
        ix2 = (rs > rs_sat * (1 + ETOL)) & st1_prev & (prev.rs > rs_sat0 * (1 - ETOL))
        sg[ix2] = 0.0
        gas |= ix2
        ix = sg < 0
        sw[ix] /= 1.0 - sg[ix]
        so[ix] /= 1.0 - sg[ix]
        sg[ix] = 0.0
        ix = so < 0
        sw[ix] /= 1.0 - so[ix]
        sg[ix] /= 1.0 - so[ix]
        so[ix] = 0.0
        ix = sw < 0
        so[ix] /= 1.0 - sw[ix]
        sg[ix] /= 1.0 - sw[ix]
        sw[ix] = 0.0
        rs[gas] = rs_sat[gas]
        rs[~gas] = np.minimum(rs_sat[~gas], rs[~gas])
        st.rs = np.maximum(rs, 0.0)
        st.status = 1 + 2 * gas.astype(np.int64)
        s = np.column_stack([sw, so, sg])
        st.s = s / s.sum(axis=1, keepdims=True)

    def converged(self, eqs, dt, st: BOState):
        f = self.fluid
        res = [value(e) for e in eqs]
        p = st.pressure
        b = [value(f.b_w(p)), value(f.b_o(p, st.rs, st.status == 3)), value(f.b_g(p))]
        b_avg = np.array([np.mean(1.0 / x) for x in b])
        pv = self.pv_ref * value(f.pv_mult(p))
        cnv = max(b_avg[i] * dt * np.abs(res[i] / pv).max() for i in range(3))
        mb = max(b_avg[i] * dt * abs(res[i].sum()) / pv.sum() for i in range(3))
        well_ok = ctrl_ok = True
        if self.wells:   # (the reference takes max over the wells, so it needs one)
            q_ref = max(abs(w.val) for w in self.wells)
            well_ok = all(np.abs(r).max() <= self.tol_well * q_ref for r in res[3:6])
            ctrl_scale = np.array([1e5 if w.type == "bhp" else abs(w.val) for w in self.wells])
            ctrl_ok = bool(np.all(np.abs(res[6]) <= self.tol_well * ctrl_scale))
        return bool(cnv <= self.tol_cnv and mb <= self.tol_mb and well_ok and ctrl_ok), (cnv, mb)

    def step(self, state: BOState, dt: float):
        st0, st = state, state.copy()
        for it in range(self.max_iterations):
            r, J, eqs = self.residual_and_jacobian(st0, st, dt)
            if self.converged(eqs, dt, st)[0]:
                return st, True, it
            st = self.update_state(st, spla.spsolve(J.tocsc(), -r))
        r, J, eqs = self.residual_and_jacobian(st0, st, dt)
        return st, self.converged(eqs, dt, st)[0], self.max_iterations


def _sum(x):
    if isinstance(x, ADI):
        import scipy.sparse as sps
        return ADI([x.val.sum()], [(sps.csr_array(np.ones((1, x.val.size))) @ j).tocsr() for j in x.jac])
    return np.array([np.sum(x)])


def _vcat(items, like=None):
    items = list(items)
    if not items:
        import scipy.sparse as sps
        if isinstance(like, ADI):
            return ADI(np.zeros(0), [sps.csr_array((0, j.shape[1])) for j in like.jac])
        return np.zeros(0)
    if any(isinstance(i, ADI) for i in items):
        import scipy.sparse as sps
        n = sum(len(value(i)) for i in items)
        ref = next(i for i in items if isinstance(i, ADI))
        val = np.concatenate([np.atleast_1d(value(i)) for i in items])
        jac = [sps.vstack([i.jac[b] if isinstance(i, ADI) else sps.csr_array((len(value(i)), ref.jac[b].shape[1]))
                           for i in items], format="csr") for b in range(len(ref.jac))]
        return ADI(val, jac)
    return np.concatenate([np.atleast_1d(i) for i in items])


def simulate_schedule(state0, model: BOModel, timesteps, max_cuts=6):
    state = state0
    states, its_per_step = [], np.zeros(len(timesteps), dtype=np.int64)
    for i, dt_report in enumerate(timesteps):
        t_local, its, dt, cuts = 0.0, 0, float(dt_report), 0
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


def synthetic_deck(nrows: int = 20, p_lo=50e5, p_hi=400e5):
    """The synthetic monotone PVT of the black-oil workload: 20 saturated
    rows over 50-400 bar with Bo 1.05 -> 1.35, Rs 0 -> 150 Sm3/Sm3,
    mu_o 2 -> 0.8 cP, each with one undersaturated row at +50 bar (Bo down
    1 %, mu_o up 5 %); Bg from Z 0.95 -> 0.85 at 373 K; Corey-2 SWOF/SGOF
    with connate water 0.2; PVTW and ROCK compressibility."""
    p = np.linspace(p_lo, p_hi, nrows)
    t = (p - p_lo) / (p_hi - p_lo)
    rs = 150.0 * t
    bo = 1.05 + 0.30 * t
    muo = (2.0 - 1.2 * t) * 1e-3
    pvto = []
    for k in range(nrows):
        rows = [[p[k], bo[k], muo[k]], [p[k] + 50e5, bo[k] * 0.99, muo[k] * 1.05]]
        pvto.append((rs[k], rows))
    Z = 0.95 - 0.10 * t
    Tk, p_sc, T_sc = 373.15, 1.01325e5, 288.71
    Bg = Z * Tk * p_sc / (p * T_sc)
    mug = (0.012 + 0.012 * t) * 1e-3
    pvdg = np.column_stack([p, Bg, mug])
    sw = np.linspace(0.2, 1.0, 9)
    swof = np.column_stack([sw, ((sw - 0.2) / 0.8) ** 2, ((1.0 - sw) / 0.8) ** 2, np.zeros_like(sw)])
    sg = np.linspace(0.0, 0.8, 9)
    sgof = np.column_stack([sg, (sg / 0.8) ** 2, ((0.8 - sg) / 0.8) ** 2, np.zeros_like(sg)])
    return dict(rhoS=[1000.0, 850.0, 0.9],
                pvtw=dict(p_ref=200e5, b_ref=1.01, c=4e-10, mu_ref=0.5e-3, viscosibility=0.0),
                rock_cp=dict(p_ref=200e5, c=5e-10), pvdg=pvdg, pvto=pvto, swof=swof, sgof=sgof)
