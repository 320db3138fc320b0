"""GenericBlackOilModelAD in its oil-water form, with the FacilityModel well
equations (the oracle: test infrastructure only, never the product path).

The generic model's reservoir equations are the hand-rolled model's (the
reference's own proof, generic_black_oil_model.py:1026-1062, reproduces
TwoPhaseOilWaterModelAD to 1e-11), so the cell terms are oracle/ow.py's.
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
from .bo import _vcat
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
