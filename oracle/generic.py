"""GenericBlackOilModelAD with the FacilityModel well equations, in its
oil-water form (simple fluid) and its three-phase live-oil form (deck fluid,
disgas) (the oracle: test infrastructure only, never the product path).

The generic model's reservoir equations are the hand-rolled models' (the
reference's own proofs, generic_black_oil_model.py:1026-1062 and
:1116-1190, reproduce TwoPhaseOilWaterModelAD to 1e-11 and
ThreePhaseBlackOilModelAD on SPE 1 to 1e-11), so the cell terms are
oracle/ow.py's and oracle/bo.py's.
What this module restates is the facility layer and the generic model's
step protocol:
  FacilityModel.well_equations              facility_model.py:2197-2292
  FacilityModel.connection_pressure_drop    facility_model.py:2294-2389
  FacilityModel._single_well                facility_model.py:2434-2549
  FacilityModel._control_equation           facility_model.py:2589-2658
  FacilityModel.update_switched_controls    facility_model.py:1103-1143
  FacilityModel._violated_limit             facility_model.py:1790-1817
  GenericBlackOilModelAD.get_equations      generic_black_oil_model.py:472-526
  GenericBlackOilModelAD.after_convergence  generic_black_oil_model.py:569-601
  GenericBlackOilModelAD.get_convergence    generic_black_oil_model.py:604-681
  well_control_residual                     ad_core/utils/convergence.py:72-112
  update_state_black_oil_generic            ad_blackoil/utils/update_state_black_oil_generic.py:10-104
with the reference's defaults: surface-volume basis, cross-flow allowed.
VFP/THP and reservoir-volume ("volume") controls and multisegment wells
are not restated.
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import scipy.sparse as sp

from .adi import ADI, ad_matmul, ad_sum, value
from .bo import BOModel, BOState, _sum, _vcat
from .ow import OWModel, OWState

RATE_TYPES = ("rate", "vrat", "resv_history")


@dataclass
class GenericOWState(OWState):
    """OWState plus the wells' perforation fluxes of the last converged step
    (MRST's wellSol.flux: reservoir volumes per perforation and phase)."""
    well_flux: Optional[List[np.ndarray]] = None

    def copy(self):
        return GenericOWState(self.pressure.copy(), self.sW.copy(), self.qWs.copy(), self.qOs.copy(),
                              self.bhp.copy(),
                              None if self.well_flux is None else [f.copy() for f in self.well_flux])


def violated_limit(w, bhp, q):
    """The first limit a well breaks other than the one it is on
    (facility_model.py:1790-1817).  Rates are signed (production < 0), so a
    producer breaks a rate limit by going below it."""
    lims = w.lims or {}
    qt = float(np.sum(q))
    if w.sign > 0:
        order = [("bhp", bhp > lims.get("bhp", np.inf)), ("rate", qt > lims.get("rate", np.inf))]
    else:
        qw = q[0]
        qo = q[1] if len(q) > 1 else 0.0
        qg = q[2] if len(q) > 2 else 0.0
        order = [("bhp", bhp < lims.get("bhp", -np.inf)), ("orat", qo < lims.get("orat", -np.inf)),
                 ("lrat", qw + qo < lims.get("lrat", -np.inf)), ("grat", qg < lims.get("grat", -np.inf)),
                 ("wrat", qw < lims.get("wrat", -np.inf)), ("vrat", qt > -lims.get("vrat", np.inf))]
    for mode, broken in order:
        if mode != w.type and broken and mode in lims:
            return mode
    return None


def control_equation(w, q_s, bhp, control_density=None):
    """The control row of one well (facility_model.py:2589-2658): a bhp
    target (or the total rate when a flowing well runs the wrong way), a
    zero rate target closing on the total rate, otherwise the named rate."""
    qt = q_s[0]
    for q in q_s[1:]:
        qt = qt + q
    if w.type == "bhp":
        qv = float(value(qt)[0])
        if (not getattr(w, "allow_sign_change", False) and w.sign != 0 and abs(qv) > 1e-8
                and np.sign(qv) != w.sign):
            return qt
        return bhp - w.val
    if float(w.val) == 0.0:
        return qt
    if w.type in RATE_TYPES:
        return qt - w.val
    if w.type == "resv":
        rho = np.ones(len(q_s)) if control_density is None else np.asarray(control_density, float)
        resv = q_s[0] * rho[0]
        for i in range(1, len(q_s)):
            resv = resv + q_s[i] * rho[i]
        return resv - w.val
    if w.type == "wrat":
        return q_s[0] - w.val
    if w.type == "orat":
        return q_s[1] - w.val
    if w.type == "grat":
        return q_s[2] - w.val
    if w.type == "lrat":
        return q_s[0] + q_s[1] - w.val
    raise NotImplementedError(f"control '{w.type}' is not restated")


def column_pressure_drop(dz_perf, wi, rho_c, mob_c, compi, sign, flux, gravity):
    """The wellbore column from the bhp datum to each perforation
    (facility_model.py:2294-2389): segment k carries the streams of
    perforations k..end, so its density is their summed mass flux over their
    summed volume flux, from last step's perforation fluxes (a producer with
    history), WI x mobility (a producer without), or compi (an injector)."""
    dz_perf = np.asarray(dz_perf, float).ravel()
    n = dz_perf.size
    if not np.any(dz_perf) or gravity == 0.0:
        return np.zeros(n)
    if sign < 1 and flux is not None and np.any(flux):
        qv = np.asarray(flux, float).reshape(n, -1)
    elif sign == -1:
        qv = wi[:, None] * mob_c
    else:
        qv = wi[:, None] * np.asarray(compi, float)[None, :rho_c.shape[1]]
    q_mass = np.sum(qv * rho_c, axis=1)
    q_vol = np.sum(qv, axis=1)
    if abs(q_mass.sum()) < 1e-20:
        qa = wi[:, None] * mob_c
        rho_mix = np.sum(qa * rho_c, axis=1) / np.maximum(np.sum(qa, axis=1), 1e-30)
    else:
        m_down = np.abs(np.cumsum(q_mass[::-1])[::-1])
        v_down = np.abs(np.cumsum(q_vol[::-1])[::-1])
        with np.errstate(divide="ignore", invalid="ignore"):
            rho_mix = m_down / v_down
        rho_mix[~np.isfinite(rho_mix)] = 0.0
        for i in range(n - 2, -1, -1):
            if rho_mix[i] < 1e-8:
                rho_mix[i] = rho_mix[i + 1]
        bad = (rho_mix == 0.0) | ~np.isfinite(rho_mix)
        if bad.all():
            return np.zeros(n)
        rho_mix[bad] = rho_mix[~bad].mean()
    dz_seg = np.diff(np.concatenate([[0.0], dz_perf]))
    return np.cumsum(gravity * rho_mix * dz_seg)


class GenericOWModel(OWModel):
    """GenericBlackOilModelAD(g, rock, fluid, wells) for a two-phase simple
    fluid: the oracle OW cell equations with the facility wells."""

    def __init__(self, *args, allow_cross_flow=True, **kw):
        super().__init__(*args, **kw)
        self.allow_cross_flow = allow_cross_flow
        self.within_limits = True
        self.last_flux = None
        self._rho_scale = np.ones(2)
        self.switches = []              # (well index, old type, new type) in order

    def initial_state(self, p0, sw0) -> GenericOWState:
        st = super().initial_state(p0, sw0)
        return GenericOWState(st.pressure, st.sW, st.qWs, st.qOs, st.bhp, None)

    # -- per-step quantities frozen at state0 --------------------------------
    def compute_cdp_per_well(self, st0: GenericOWState):
        """The facility's column drops and rhoScale, both at the previous
        state (facility_model.py:2212-2276)."""
        f = self.fluid
        p0 = st0.pressure
        b0 = [value(f.b(0, p0)), value(f.b(1, p0))]
        self._rho_scale = np.array([f.rhoS[i] / np.mean(f.rhoS[i] * b0[i]) for i in range(2)])
        mob0 = [value(f.kr(0, st0.sW)) / f.mu[0], value(f.kr(1, 1.0 - st0.sW)) / f.mu[1]]
        out = []
        for k, w in enumerate(self.wells):
            wc = w.cells
            rho_c = np.column_stack([f.rhoS[i] * b0[i][wc] for i in range(2)])
            mob_c = np.column_stack([mob0[i][wc] for i in range(2)])
            flux = None if st0.well_flux is None else st0.well_flux[k]
            out.append(column_pressure_drop(w.dZ, np.asarray(w.WI, float), rho_c, mob_c, w.compi, w.sign,
                                            flux, self.gravity))
        return out

    # -- facility well equations --------------------------------------------
    def _single_well(self, w, p, mob, b, q_s, bhp_k, cdp):
        """_single_well (facility_model.py:2434-2549), two phases."""
        wc, tw = w.cells, np.asarray(w.WI, float)
        compi = [float(c) for c in w.compi[:2]]
        dd = -(bhp_k + cdp) + p[wc]
        if not getattr(w, "status", True):
            zero = 0.0 * dd
            ref = ad_sum(p[wc]) / max(wc.size, 1)
            return [zero, zero], [q_s[0], q_s[1]], bhp_k - ref, [np.zeros(wc.size)] * 2
        inj = value(dd) < 0
        prod = ~inj
        mob_c = [m[wc] for m in mob]
        b_c = [x[wc] for x in b]
        cq_p = cq_ps = None
        mix = compi
        if prod.any():
            con = prod * tw
            cq_p = [-con * m * dd for m in mob_c]
            cq_ps = [b_c[i] * cq_p[i] for i in range(2)]
            wbq = [q_s[i] * float(value(q_s[i])[0] > 0) - ad_sum(cq_ps[i]) for i in range(2)]
            wbqt = wbq[0] + wbq[1]
            if self.allow_cross_flow and value(wbqt)[0] != 0.0:
                mix = [q / wbqt for q in wbq]
        if inj.any():
            cqt_i = -(inj * tw) * ((mob_c[0] + mob_c[1]) * dd)
            cqt_is = cqt_i / (mix[0] / b_c[0] + mix[1] / b_c[1])
            cq_s = [(cq_ps[i] if cq_ps is not None else 0.0) + mix[i] * cqt_is for i in range(2)]
        else:
            cq_s = cq_ps
        cq_vol = []
        for i in range(2):
            v = np.zeros(wc.size)
            if inj.any():
                v = v + inj * (value(cqt_i) * compi[i])
            if prod.any():
                v = v + prod * value(cq_p[i])
            cq_vol.append(v)
        weqs = [(q_s[i] - ad_sum(cq_s[i])) * self._rho_scale[i] for i in range(2)]
        return cq_s, weqs, control_equation(w, q_s, bhp_k), cq_vol

    def well_contributions(self, p, mob_w, mob_o, b_w, b_o, qws, qos, bhp, cdp_per_well):
        """get_equations' well part (generic_black_oil_model.py:496-525): the
        cell adjustments, the rate-closure rows, the control rows."""
        self.within_limits = True
        nc = self.nc
        adj = [None, None]
        eqw, eqo, ctrl, flux = [], [], [], []
        for k, w in enumerate(self.wells):
            kk = np.array([k])
            q_s = [qws[kk], qos[kk]]
            cq_s, weqs, c, cq_vol = self._single_well(w, p, [mob_w, mob_o], [b_w, b_o], q_s, bhp[kk],
                                                      np.asarray(cdp_per_well[k], float))
            S = sp.csr_array((np.ones(w.cells.size), (w.cells, np.arange(w.cells.size))), shape=(nc, w.cells.size))
            for i in range(2):
                contrib = ad_matmul(S, cq_s[i])
                adj[i] = contrib if adj[i] is None else adj[i] + contrib
            eqw.append(weqs[0])
            eqo.append(weqs[1])
            ctrl.append(c)
            flux.append(np.column_stack(cq_vol))
        self.last_flux = flux
        if adj[0] is None:
            adj = [np.zeros(nc), np.zeros(nc)]
        like = p if isinstance(p, ADI) else None
        return adj[0], adj[1], _vcat(eqw, like), _vcat(eqo, like), _vcat(ctrl, like), None, None

    # -- convergence ----------------------------------------------------------
    def converged(self, eqs, dt, st):
        """get_convergence (generic_black_oil_model.py:640-681): CNV and MB
        with B_avg = mean(1/b) and the current pore volume, the scaled well
        and control residuals, and the within-limits flag."""
        water, oil, eq_w, eq_o, ctrl = (value(e) for e in eqs)
        f = self.fluid
        pvm = f.pv_mult(st.pressure)
        pv = self.pv * (1.0 if pvm is None else value(pvm))
        b_avg = [np.mean(1.0 / value(f.b(i, st.pressure))) for i in range(2)]
        cnv = max(b_avg[0] * dt * np.abs(water / pv).max(), b_avg[1] * dt * np.abs(oil / pv).max())
        mb = max(dt * abs(b_avg[0] * water.sum()) / pv.sum(), dt * abs(b_avg[1] * oil.sum()) / pv.sum())
        well_res, ctrl_res = well_control_residual([eq_w, eq_o], ctrl, self.wells)
        ok = (cnv <= self.tol_cnv and mb <= self.tol_mb and well_res <= self.tol_well
              and ctrl_res <= self.tol_well and self.within_limits)
        self.last_residuals = {"CNV": cnv, "MB": mb, "well": well_res, "control": ctrl_res}
        return bool(ok), (cnv, mb)

    # -- update -----------------------------------------------------------------
    def apply_update(self, st, upd):
        """update_state_black_oil_generic's oil-water branch, then the
        facility's control switching."""
        super().apply_update(st, upd)
        self.update_switched_controls(st)

    def update_switched_controls(self, st):
        ok = True
        for k, w in enumerate(self.wells):
            if not getattr(w, "status", True) or not w.lims:
                continue
            mode = violated_limit(w, float(st.bhp[k]), np.array([st.qWs[k], st.qOs[k]]))
            if mode is None:
                continue
            ok = False
            self.switches.append((k, w.type, mode))
            w.type = mode
            w.val = float(w.lims[mode])
        self.within_limits = ok
        return ok

    def after_convergence(self, st, st0, dt):
        """The perforation fluxes of the converged step go with its state
        (generic_black_oil_model.py:569-601)."""
        st.well_flux = [q.copy() for q in self.last_flux] if self.last_flux is not None else None
        return st


def well_control_residual(well_residuals, control_residual, wells):
    """convergence.py:72-112: rate rows over the largest nonzero scheduled
    rate (1 when there is none), control rows over 1 bar for a bhp control,
    over the target for a nonzero rate control and over the reference rate
    (at least 1) for a zero one."""
    if not len(wells):
        return 0.0, 0.0
    positive = [abs(w.val) for w in wells if w.type != "bhp" and abs(w.val) > 1e-12]
    q_ref = max(positive) if positive else 1.0
    well = max(np.abs(np.asarray(r, float)).max() for r in well_residuals) / q_ref
    scale = np.array([1e5 if w.type == "bhp" else (abs(w.val) if abs(w.val) > 1e-12 else max(q_ref, 1.0))
                      for w in wells])
    return float(well), float(np.max(np.abs(np.asarray(control_residual, float)) / scale))


# ---------------------------------------------------------------------------
# Three-phase live oil (disgas): GenericBlackOilModelAD(g, rock, deck fluid,
# wells, disgas=True) with the facility wells.
# ---------------------------------------------------------------------------


@dataclass
class GenericBOState(BOState):
    """BOState plus the wells' perforation fluxes of the last converged
    step (MRST's wellSol.flux), one (nperf, 3) array per well."""
    well_flux: Optional[List[np.ndarray]] = None

    def copy(self):
        b = BOState.copy(self)
        return GenericBOState(b.pressure, b.s, b.rs, b.status, b.qWs, b.qOs, b.qGs, b.bhp,
                              None if self.well_flux is None else [f.copy() for f in self.well_flux])


class GenericBOModel(BOModel):
    """GenericBlackOilModelAD(g, rock, deck fluid, wells, disgas=True): the
    three-phase cell equations of oracle/bo.py (the reference's own proof,
    generic_black_oil_model.py:1116-1190, reproduces ThreePhaseBlackOilModelAD
    on SPE 1 to 1e-11) with what differs from the hand-rolled model:
      the facility wells (_single_well facility_model.py:2434-2549 with
        dissolved gas carried by the oil stream :2466-2475 and the mixture's
        volume ratio _volume_ratio :2391-2432), the wellbore column
        (connection_pressure_drop :2294-2389) from reservoir densities with
        dissolved gas, rhoScale from the whole grid's densities at state0
        (:2239-2263), every control (_control_equation :2589-2658), shut wells,
        limit switching (update_switched_controls :1103-1143);
      update_state_black_oil_generic (update_state_black_oil_generic.py:30-99):
        pressure change capped at dp_max_rel |p|, rs change at drs_max, the
        saturation changes scaled together so the largest is ds_max, then the
        flash (compute_flash_black_oil.py:22-140, oracle/bo.py's flash);
      get_convergence (generic_black_oil_model.py:640-681): CNV/MB as the
        hand-rolled model's (B_avg = mean(1/b), current pore volume) with the
        wells scaled by convergence.py:72-112 and within_limits."""

    def __init__(self, g, rock, fluid, wells, N, T, gravity=0.0, dp_max_rel=0.2, drs_max=np.inf,
                 allow_cross_flow=True, **kw):
        super().__init__(g, rock, fluid, wells, N, T, gravity=gravity, **kw)
        self.gravity = float(gravity)
        self.dp_max_rel, self.drs_max = float(dp_max_rel), float(drs_max)
        self.allow_cross_flow = allow_cross_flow
        self.within_limits = True
        self.last_flux = None
        self.switches = []
        self._rho_scale = np.ones(3)
        self._cdp = None
        self._flux = None

    def initial_state(self, *args, **kw) -> GenericBOState:
        b = super().initial_state(*args, **kw)
        return GenericBOState(b.pressure, b.s, b.rs, b.status, b.qWs, b.qOs, b.qGs, b.bhp, None)

    # -- reservoir densities and mobilities at a state (values) -------------
    def densities(self, st):
        """(rho (nc, 3), mobility (nc, 3)): b_w rhoS_w, b_o (rhoS_o + rs
        rhoS_g), b_g rhoS_g; the saturated cells' rs is rs_sat(p) and an
        undersaturated cell has no free gas, as in the assembly."""
        f = self.fluid
        p = st.pressure
        st1, st3 = st.status == 1, st.status == 3
        rs = np.where(st1, st.rs, value(f.rs_sat(p)))
        sw = st.s[:, 0]
        sg = np.where(st3, st.s[:, 2], 0.0)
        so = 1.0 - sw - sg
        b = [value(f.b_w(p)), value(f.b_o(p, rs, st3)), value(f.b_g(p))]
        rho = np.column_stack([b[0] * f.rhoS[0], b[1] * (f.rhoS[1] + rs * f.rhoS[2]), b[2] * f.rhoS[2]])
        mu = [value(f.mu_w(p)), value(f.mu_o(p, rs, st3)), value(f.mu_g(p))]
        kr = [value(f.kr_w(sw)), value(f.kr_o(sw, so, sg)), value(f.kr_g(sg))]
        return rho, np.column_stack([kr[i] / mu[i] for i in range(3)])

    def compute_cdp_per_well(self, st0):
        """rhoScale and the column drops, both at state0 (facility_model.py:2212-2276)."""
        rho, mob = self.densities(st0)
        mean = rho.mean(axis=0)
        self._rho_scale = np.array([self.fluid.rhoS[i] / mean[i] if np.isfinite(mean[i]) and mean[i] > 0 else 1.0
                                    for i in range(3)])
        out = []
        for k, w in enumerate(self.wells):
            wc = w.cells
            flux = None if getattr(st0, "well_flux", None) is None else st0.well_flux[k]
            out.append(column_pressure_drop(w.dZ, np.asarray(w.WI, float), rho[wc], mob[wc], w.compi, w.sign,
                                            flux, self.gravity))
        return out

    def residual_and_jacobian(self, st0, st, dt, cdp=None):
        self.within_limits = True
        self._cdp = cdp if cdp is not None else [np.zeros(w.cells.size) for w in self.wells]
        self._flux = [None] * len(self.wells)
        out = super().residual_and_jacobian(st0, st, dt)
        self.last_flux = list(self._flux)
        return out

    def _well(self, k, p, mob, b, rs, q_s, bhp_k):
        """_single_well for three phases with dissolved gas."""
        w = self.wells[k]
        wc, tw = w.cells, np.asarray(w.WI, float)
        compi = [float(c) for c in np.asarray(w.compi, float)[:3]]
        dd = -(bhp_k + np.asarray(self._cdp[k], float)) + p[wc]
        if not getattr(w, "status", True):
            zero = 0.0 * dd
            self._flux[k] = np.zeros((wc.size, 3))
            ref = ad_sum(p[wc]) / max(wc.size, 1)
            return [zero, zero, zero], [q_s[0], q_s[1], q_s[2]], bhp_k - ref
        inj = value(dd) < 0
        prod = ~inj
        mob_c = [m[wc] for m in mob]
        b_c = [x[wc] for x in b]
        rs_c = rs[wc]
        any_inj, any_prod = bool(inj.any()), bool(prod.any())
        cq_p = cq_ps = None
        mix = list(compi)
        if any_prod:
            con = prod * tw
            cq_p = [-con * m * dd for m in mob_c]
            cq_ps = [b_c[i] * cq_p[i] for i in range(3)]
            cq_ps[2] = cq_ps[2] + rs_c * (b_c[1] * cq_p[1])
            wbq = [q_s[i] * float(value(q_s[i])[0] > 0) - _sum(cq_ps[i]) for i in range(3)]
            wbqt = wbq[0] + wbq[1] + wbq[2]
            if self.allow_cross_flow and value(wbqt)[0] != 0.0:
                mix = [q / wbqt for q in wbq]
        if any_inj:
            mt = mob_c[0] + mob_c[1] + mob_c[2]
            cqt_i = -(inj * tw) * (mt * dd)
            vol_rat = mix[0] / b_c[0] + mix[1] / b_c[1] + (mix[2] - rs_c * mix[1]) / b_c[2]
            cqt_is = cqt_i / vol_rat
            cq_s = [(cq_ps[i] if any_prod else 0.0) + mix[i] * cqt_is for i in range(3)]
        else:
            cq_s = cq_ps
        vol = np.zeros((wc.size, 3))
        for i in range(3):
            if any_inj:
                vol[:, i] += inj * (value(cqt_i) * compi[i])
            if any_prod:
                vol[:, i] += prod * value(cq_p[i])
        self._flux[k] = vol
        weqs = [(q_s[i] - _sum(cq_s[i])) * self._rho_scale[i] for i in range(3)]
        return cq_s, weqs, control_equation(w, list(q_s), bhp_k)

    # -- convergence ----------------------------------------------------------
    def converged(self, eqs, dt, st):
        f = self.fluid
        res = [value(e) for e in eqs]
        p = st.pressure
        b = [value(f.b_w(p)), value(f.b_o(p, st.rs, st.status == 3)), value(f.b_g(p))]
        b_avg = np.array([np.mean(1.0 / x) for x in b])
        pv = self.pv_ref * value(f.pv_mult(p))
        cnv = max(b_avg[i] * dt * np.abs(res[i] / pv).max() for i in range(3))
        mb = max(b_avg[i] * dt * abs(res[i].sum()) / pv.sum() for i in range(3))
        well_res, ctrl_res = well_control_residual(res[3:6], res[6], self.wells)
        ok = (cnv <= self.tol_cnv and mb <= self.tol_mb and well_res <= self.tol_well and ctrl_res <= self.tol_well
              and self.within_limits)
        self.last_residuals = {"CNV": cnv, "MB": mb, "well": well_res, "control": ctrl_res}
        return bool(ok), (cnv, mb)

    def solve_increment(self, r, J, st):
        import scipy.sparse.linalg as spla
        return spla.spsolve(J.tocsc(), -r)

    # -- update ----------------------------------------------------------------
    def apply_update(self, st, upd):
        """update_state_black_oil_generic with the flash, in place, then the
        facility's control switching."""
        nc, nw = self.nc, len(self.wells)
        prev = st.copy()
        st1, st3 = prev.status == 1, prev.status == 3
        dp = upd[:nc]
        cap = self.dp_max_rel * np.abs(prev.pressure)
        st.pressure = prev.pressure + np.clip(dp, -cap, cap)
        dsw, dx_hc = upd[nc:2 * nc], upd[2 * nc:3 * nc]
        dsg = np.where(st3, dx_hc, 0.0)
        drs = np.clip(np.where(st1, dx_hc, 0.0), -self.drs_max, self.drs_max)
        st.rs = np.maximum(prev.rs + drs, 0.0)
        ds = np.column_stack([dsw, -(dsg + dsw), dsg])
        largest = np.abs(ds).max(axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            scale = np.minimum(np.divide(self.ds_max, largest, out=np.ones_like(largest), where=largest > 0), 1.0)
        st.s = prev.s + ds * scale[:, None]
        off = 3 * nc
        st.qWs = prev.qWs + upd[off:off + nw]
        st.qOs = prev.qOs + upd[off + nw:off + 2 * nw]
        st.qGs = prev.qGs + upd[off + 2 * nw:off + 3 * nw]
        st.bhp = prev.bhp + upd[off + 3 * nw:off + 4 * nw]
        self.flash(st, prev, st1)
        self.update_switched_controls(st)

    def update_switched_controls(self, st):
        ok = True
        for k, w in enumerate(self.wells):
            if not getattr(w, "status", True) or not w.lims:
                continue
            mode = violated_limit(w, float(st.bhp[k]), np.array([st.qWs[k], st.qOs[k], st.qGs[k]]))
            if mode is None:
                continue
            ok = False
            self.switches.append((k, w.type, mode))
            w.type = mode
            w.val = float(w.lims[mode])
        self.within_limits = ok
        return ok

    def after_convergence(self, st, st0, dt):
        st.well_flux = [q.copy() for q in self.last_flux] if self.last_flux is not None else None
        return st
