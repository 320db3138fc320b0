"""GenericBlackOilModelAD on the MI355X with the FacilityModel well
equations: the two-phase (oil-water) simple-fluid form, and the three-phase
live-oil form (a deck fluid with dissolved gas) at the end of this module.

Reference: pymrst/modules/ad_blackoil/models/generic_black_oil_model.py
(__init__ :99-196, get_equations :472-526, after_convergence :569-601,
get_convergence :604-681, update_state :683-684) with
pymrst/modules/ad_core/models/facilities/facility_model.py
(well_equations :2197-2292, connection_pressure_drop :2294-2389,
_single_well :2434-2549, _control_equation :2589-2658,
update_switched_controls :1103-1143, _violated_limit :1790-1817) and
pymrst/modules/ad_core/utils/convergence.py:72-112.

The reservoir equations are the hand-rolled oil-water model's (the
reference's own proof reproduces TwoPhaseOilWaterModelAD to 1e-11), so the
device assembly is pmx_ow's; the facility differs from the hand-rolled
wells in four places, all carried to the device through pmx_ow_set_controls
and the per-step column drops:
* every control the facility closes a well with (bhp, total / phase /
  liquid rate, reservoir rate with control densities, a zero target's total
  rate, the bhp well's fall back to the total rate when it flows the wrong
  way) as a linear form of the well's unknowns, chosen on the host before
  each assembly from the well's current rates;
* a well that breaks a limit switches to it (update_switched_controls,
  after every Newton update), and shut wells;
* the rate-closure rows scaled by rhoScale = rhoS / mean(rho) at the step's
  start;
* the wellbore column from last step's perforation fluxes (cumulated from
  the toe), frozen over the step, handed on by after_convergence.
Convergence is the generic model's: B_avg = mean(1/b), the current pore
volume, the well and control rows scaled by convergence.py's rule.

Not on the device: vaporised oil, three-phase dead oil, reservoir-volume
("volume") and VFP/THP controls, multisegment wells, boundary conditions,
sources and aquifers (each raises).  The linear solver is the device
CPR-FGMRES.  Drive it with pymrst_amd.nonlinear.NonLinearSolver
(solve_ministep / solve_timestep / simulate_schedule).
"""
from dataclasses import dataclass
from typing import List, Optional, Sequence

import ctypes
import numpy as np
import torch

from . import _lib
from .bo import BOState, DeckBlackOilFluid, DeviceBOState, ThreePhaseBlackOilModelAD
from .grid import Grid, Rock, Well
from .ow import CPRSolver, DeviceOWState, OWState, SimpleADIFluid, TwoPhaseOilWaterModelAD

_RATE_TYPES = ("rate", "vrat", "resv_history")
_KNOWN = ("bhp",) + _RATE_TYPES + ("resv", "wrat", "orat", "lrat")


@dataclass
class GenericState(OWState):
    """Host state with the wells' perforation fluxes of the last converged
    step (MRST's wellSol.flux), one (nperf, 2) array per well."""
    well_flux: Optional[List[np.ndarray]] = None

    def copy(self):
        base = super().copy()
        return GenericState(base.pressure, base.sW, base.qWs, base.qOs, base.bhp,
                            None if self.well_flux is None else [f.copy() for f in self.well_flux])


@dataclass
class DeviceGenericState(DeviceOWState):
    well_flux: Optional[List[np.ndarray]] = None

    def copy(self):
        base = super().copy()
        return DeviceGenericState(base.p, base.sw, base.qws, base.qos, base.bhp,
                                  None if self.well_flux is None else [f.copy() for f in self.well_flux])

    def to_host(self) -> GenericState:
        h = super().to_host()
        return GenericState(h.pressure, h.sW, h.qWs, h.qOs, h.bhp,
                            None if self.well_flux is None else [f.copy() for f in self.well_flux])

    @staticmethod
    def from_host(st, device) -> "DeviceGenericState":
        d = DeviceOWState.from_host(st, device)
        flux = getattr(st, "well_flux", None)
        return DeviceGenericState(d.p, d.sw, d.qws, d.qos, d.bhp,
                                  None if flux is None else [np.array(f, float) for f in flux])


def limit_broken(w: Well, bhp: float, q: Sequence[float]) -> Optional[str]:
    """The limit a well breaks first, other than the control it is on
    (_violated_limit): an injector's bhp ceiling and total-rate cap; a
    producer's bhp floor and its oil, liquid, gas, water and volume-rate
    floors, in that order.  Production rates are negative."""
    lims = getattr(w, "lims", None) or {}
    total = float(np.sum(q))
    if w.sign > 0:
        tests = (("bhp", lambda v: bhp > v), ("rate", lambda v: total > v))
    else:
        qw, qo = float(q[0]), float(q[1]) if len(q) > 1 else 0.0
        qg = float(q[2]) if len(q) > 2 else 0.0
        tests = (("bhp", lambda v: bhp < v), ("orat", lambda v: qo < v), ("lrat", lambda v: qw + qo < v),
                 ("grat", lambda v: qg < v), ("wrat", lambda v: qw < v), ("vrat", lambda v: total > -v))
    for mode, broken in tests:
        if mode in lims and mode != w.type and broken(float(lims[mode])):
            return mode
    return None


def control_form(w: Well, qw: float, qo: float):
    """(coefficients on qWs, qOs, bhp; target) of a well's control row."""
    if w.type not in _KNOWN:
        raise NotImplementedError(f"well control '{w.type}' is not available on the device")
    if w.type == "bhp":
        qt = qw + qo
        if (not getattr(w, "allow_sign_change", False) and w.sign != 0 and abs(qt) > 1e-8
                and np.sign(qt) != w.sign):
            return (1.0, 1.0, 0.0), 0.0          # flowing the wrong way: hold the total rate at zero
        return (0.0, 0.0, 1.0), float(w.val)
    if float(w.val) == 0.0:
        return (1.0, 1.0, 0.0), 0.0              # a zero target closes on the total rate
    if w.type in _RATE_TYPES or w.type == "lrat":
        return (1.0, 1.0, 0.0), float(w.val)
    if w.type == "wrat":
        return (1.0, 0.0, 0.0), float(w.val)
    if w.type == "orat":
        return (0.0, 1.0, 0.0), float(w.val)
    rho = getattr(w, "control_density", getattr(w, "ControlDensity", None))
    rho = np.ones(2) if rho is None else np.resize(np.asarray(rho, float).ravel(), 2)
    return (float(rho[0]), float(rho[1]), 0.0), float(w.val)      # resv


def column_drops(dz, wi, rho_c, mob_c, compi, sign, flux, gravity) -> np.ndarray:
    """Pressure drop from the bhp datum to each perforation: each segment of
    the well carries the streams entering at and below it, and weighs their
    summed mass rate over their summed volume rate.  The streams are last
    step's perforation fluxes for a producer that has flowed, WI x mobility
    for one that has not, and compi for an injector; a column with no mass
    rate at all is weighed perforation by perforation with WI x mobility."""
    dz = np.asarray(dz, float).ravel()
    n = dz.size
    if gravity == 0.0 or not np.any(dz):
        return np.zeros(n)
    nph = rho_c.shape[1]
    if sign < 1 and flux is not None and np.any(flux):
        q = np.asarray(flux, float).reshape(n, -1)
    elif sign == -1:
        q = wi[:, None] * mob_c
    else:
        q = wi[:, None] * np.asarray(compi, float).ravel()[None, :nph]
    mass, vol = (q * rho_c).sum(axis=1), q.sum(axis=1)
    if abs(mass.sum()) < 1e-20:
        w_ = wi[:, None] * mob_c
        rho = (w_ * rho_c).sum(axis=1) / np.maximum(w_.sum(axis=1), 1e-30)
    else:
        below_m = np.abs(np.cumsum(mass[::-1])[::-1])
        below_v = np.abs(np.cumsum(vol[::-1])[::-1])
        with np.errstate(divide="ignore", invalid="ignore"):
            rho = below_m / below_v
        rho = np.where(np.isfinite(rho), rho, 0.0)
        for i in range(n - 2, -1, -1):          # a dead perforation takes the segment below it
            if rho[i] < 1e-8:
                rho[i] = rho[i + 1]
        good = (rho != 0.0) & np.isfinite(rho)
        if not good.any():
            return np.zeros(n)
        rho[~good] = rho[good].mean()
    return np.cumsum(gravity * rho * np.diff(np.concatenate([[0.0], dz])))


class GenericBlackOilModelAD:
    """Drop-in for GenericBlackOilModelAD(g, rock, fluid, wells, ...) with a
    two-phase simple fluid, on one GPU.  Wells may carry ``lims`` (a dict of
    limits to switch to) and ``status`` (False: shut); a switch changes the
    well's ``type`` and ``val`` in place, as the reference's facility does."""

    n_phases = 2

    def __new__(cls, g=None, rock=None, fluid=None, *args, **kw):
        # a deck (three-phase) fluid makes the live-oil form
        if cls is GenericBlackOilModelAD and isinstance(fluid, DeckBlackOilFluid):
            return super().__new__(GenericLiveOilModelAD)
        return super().__new__(cls)

    def __init__(self, g: Grid, rock: Rock, fluid: SimpleADIFluid, wells: Sequence[Well], disgas: bool = False,
                 gravity: float = 0.0, ds_max: float = 0.2, dp_max_rel: float = 0.2, tol_cnv: float = 1e-3,
                 tol_mb: float = 1e-7, tol_well: float = 1e-6, bc=None, src=None, vapoil: bool = False,
                 aquifer=None, mass_basis: bool = False, operators=None, linear_solver: Optional[CPRSolver] = None,
                 device=None, allow_cross_flow: bool = True, max_iterations: int = 25):
        if disgas or vapoil or not isinstance(fluid, SimpleADIFluid):
            raise NotImplementedError("the device generic model takes a two-phase simple fluid "
                                      "(no dissolved gas or vaporised oil)")
        if bc is not None or src is not None or aquifer is not None or operators is not None:
            raise NotImplementedError("boundary conditions, sources, aquifers and custom operators are not "
                                      "available on the device")
        if mass_basis:
            raise NotImplementedError("the device generic model assembles in surface volumes")
        if not allow_cross_flow:
            raise NotImplementedError("cross-flow is always allowed on the device (the reference's default)")
        self.wells = list(wells)
        for w in self.wells:
            control_form(w, 0.0, 0.0)             # rejects controls the device does not close
        legacy = [Well(cells=w.cells, type="bhp" if w.type == "bhp" else "rate", val=float(w.val), WI=w.WI,
                       dZ=w.dZ, compi=w.compi, sign=w.sign, name=w.name) for w in self.wells]
        self._dev = TwoPhaseOilWaterModelAD(g, rock, fluid, legacy, gravity=gravity, ds_max=ds_max,
                                            dp_max_rel=dp_max_rel, linear_solver=linear_solver, device=device,
                                            max_iterations=max_iterations)
        self.g, self.rock, self.fluid = g, rock, fluid
        self.gravity = float(gravity)
        self.tol_cnv, self.tol_mb, self.tol_well = tol_cnv, tol_mb, tol_well
        self.max_iterations = max_iterations
        self.nc, self.nw = self._dev.nc, self._dev.nw
        self.device = self._dev.device
        self.within_limits = True
        self.switches = []                          # (well, old control, new control) in order
        self._rho_scale = np.ones(2)
        self._cdp = None
        self._pmean = 0.0

    # -- reference-style helpers ---------------------------------------------
    @property
    def linear_solver(self):
        return self._dev.linear_solver

    @property
    def stats(self):
        return self._dev.stats

    def initial_state(self, p0: float, sw0: float) -> GenericState:
        st = self._dev.initial_state(p0, sw0)
        return GenericState(st.pressure, st.sW, st.qWs, st.qOs, st.bhp, None)

    def device_state(self, st) -> DeviceGenericState:
        """A device state from ours, from an OWState-like host state or from
        the reference's mapping State (pressure, sW, qWs, qOs, bhp and
        optionally well_flux)."""
        if isinstance(st, DeviceGenericState):
            return st
        if isinstance(st, DeviceOWState):
            return DeviceGenericState(st.p, st.sw, st.qws, st.qos, st.bhp, None)
        if hasattr(st, "keys"):
            st = GenericState(np.asarray(st["pressure"], float), np.asarray(st["sW"], float),
                              np.asarray(st["qWs"], float), np.asarray(st["qOs"], float),
                              np.asarray(st["bhp"], float), st.get("well_flux"))
        return DeviceGenericState.from_host(st, self.device)

    # -- per ministep: rhoScale and the column drops at state0 ----------------
    def begin_ministep(self, state0: DeviceGenericState):
        f = self.fluid
        p0 = state0.p
        for i in range(2):
            c = float(f.c[i])
            mean_b = 1.0 if c == 0.0 else float(torch.exp(c * (p0 - f.p_ref)).mean())
            self._rho_scale[i] = 1.0 / mean_b       # rhoS / mean(rhoS b)
        self._cdp = None
        if self.gravity and self.nw:
            dev = self._dev
            cells = dev.perf_cells_t
            pp = state0.p[cells].cpu().numpy()
            sp = state0.sw[cells].cpu().numpy()
            rho_c = np.column_stack([f.rhoS[i] * f.b(i, pp) for i in range(2)])
            mob_c = np.column_stack([sp ** f.n[0] / f.mu[0], (1.0 - sp) ** f.n[1] / f.mu[1]])
            out = []
            for k, w in enumerate(self.wells):
                a, b = dev.perf_ptr[k], dev.perf_ptr[k + 1]
                flux = None if state0.well_flux is None else state0.well_flux[k]
                out.append(column_drops(w.dZ, np.asarray(w.WI, float), rho_c[a:b], mob_c[a:b], w.compi, w.sign,
                                        flux, self.gravity))
            cdp = np.concatenate(out)
            dev._cdp_dev[:cdp.size].copy_(torch.as_tensor(cdp, dtype=torch.float64))
            self._cdp = dev._cdp_dev
            self.last_cdp = out

    # -- one Newton iteration: controls, assembly, convergence -----------------
    def _set_controls(self, st: DeviceGenericState):
        if not self.nw:
            return
        qw = st.qws.cpu().numpy()
        qo = st.qos.cpu().numpy()
        coef, val = np.zeros((self.nw, 3)), np.zeros(self.nw)
        for k, w in enumerate(self.wells):
            c, v = control_form(w, float(qw[k]), float(qo[k]))
            coef[k], val[k] = c, v
        status = np.array([1 if getattr(w, "status", True) else 0 for w in self.wells], dtype=np.int32)
        coef = np.ascontiguousarray(coef.ravel())
        rs = np.ascontiguousarray(self._rho_scale, dtype=np.float64)
        _lib.check(self._dev.lib.pmx_ow_set_controls(self._dev.h, _lib.ptr(coef), _lib.ptr(val), _lib.ptr(status),
                                                     _lib.ptr(rs)), "pmx_ow_set_controls")

    def linearize(self, state0: DeviceGenericState, st: DeviceGenericState, dt: float):
        """get_equations + get_convergence: (converged, {CNV, MB, well, control})."""
        self.within_limits = True
        self._set_controls(st)
        dev = self._dev
        dev.assemble_device(state0, st, dt, self._cdp)
        out8 = np.zeros(8)
        wr = np.zeros(max(3 * self.nw, 1))
        _lib.check(dev.lib.pmx_ow_convergence_current_pv(dev.h, _lib.ptr(st.p), _lib.ptr(out8), _lib.ptr(wr)),
                   "pmx_ow_convergence_current_pv")
        nc = self.nc
        b_avg = out8[4:6] / nc
        pv_sum = out8[7]
        cnv = max(b_avg[0] * dt * out8[0], b_avg[1] * dt * out8[1])
        mb = max(dt * abs(b_avg[0] * out8[2]) / pv_sum, dt * abs(b_avg[1] * out8[3]) / pv_sum)
        well_res, ctrl_res = self._well_residuals(wr[:3 * self.nw])
        self._pmean = out8[6] / nc
        ok = (cnv <= self.tol_cnv and mb <= self.tol_mb and well_res <= self.tol_well and ctrl_res <= self.tol_well
              and self.within_limits)
        return bool(ok), {"CNV": float(cnv), "MB": float(mb), "well": float(well_res), "control": float(ctrl_res)}

    def _well_residuals(self, wr: np.ndarray):
        """Rate rows over the largest nonzero scheduled rate (1 if none);
        control rows over 1 bar (bhp), the target (nonzero rate) or the
        reference rate, at least 1 (zero rate)."""
        if not self.nw:
            return 0.0, 0.0
        rates = [abs(w.val) for w in self.wells if w.type != "bhp" and abs(w.val) > 1e-12]
        q_ref = max(rates) if rates else 1.0
        well = max(np.abs(wr[0::3]).max(), np.abs(wr[1::3]).max()) / q_ref
        scale = np.array([1e5 if w.type == "bhp" else (abs(w.val) if abs(w.val) > 1e-12 else max(q_ref, 1.0))
                          for w in self.wells])
        return float(well), float(np.max(np.abs(wr[2::3]) / scale))

    # -- Newton update and control switching ----------------------------------
    def _switch_controls(self, st: DeviceGenericState):
        if not any(getattr(w, "lims", None) for w in self.wells):
            return
        bhp = st.bhp.cpu().numpy()
        q = np.column_stack([st.qws.cpu().numpy(), st.qos.cpu().numpy()])
        for k, w in enumerate(self.wells):
            if not getattr(w, "status", True) or not getattr(w, "lims", None):
                continue
            mode = limit_broken(w, float(bhp[k]), q[k])
            if mode is None:
                continue
            self.switches.append((k, w.type, mode))
            w.type, w.val = mode, float(w.lims[mode])
            self.within_limits = False

    def newton_update(self, st: DeviceGenericState, scale: float = 1.0) -> bool:
        self._dev.stats["newton"] += 1
        ok = self._dev.solve_and_update(st, self._pmean, scale)
        if ok:
            self._switch_controls(st)
        return ok

    def apply_increment(self, st: DeviceGenericState, scale: float = 1.0):
        self._dev.apply_increment(st, scale)
        self._switch_controls(st)

    @property
    def last_linear_solve(self):
        return self._dev.last_linear_solve

    def after_convergence(self, st: DeviceGenericState, state0, dt: float) -> DeviceGenericState:
        """The converged perforation fluxes go with the state, for the next
        step's wellbore column."""
        if self.nw:
            dev = self._dev
            rates = np.zeros(2 * int(dev.perf_ptr[-1]))
            _lib.check(dev.lib.pmx_ow_perf_volume_rates(dev.h, _lib.ptr(rates)), "pmx_ow_perf_volume_rates")
            st.well_flux = [rates.reshape(-1, 2)[dev.perf_ptr[k]:dev.perf_ptr[k + 1]].copy()
                            for k in range(self.nw)]
        return st


# ---------------------------------------------------------------------------
# Three-phase live oil: GenericBlackOilModelAD(g, rock, deck fluid, wells,
# disgas=True).  The reservoir equations are the hand-rolled
# ThreePhaseBlackOilModelAD's (the reference's own proof reproduces it on
# SPE 1 to 1e-11, generic_black_oil_model.py:1116-1190), so the device
# assembly is pmx_bo's; the facility reaches the device through
# pmx_bo_set_controls (control rows, shut wells, rhoScale from the whole
# grid's densities at state0 with the dissolved gas in the oil's,
# facility_model.py:2239-2263), pmx_bo_set_connection_drops (the wellbore
# column from last step's perforation fluxes and the reservoir densities at
# state0, :2294-2389) and pmx_bo_set_update_rule (update_state_black_oil_
# generic.py:30-99: pressure change capped at dp_max_rel |p|, saturation
# changes scaled together to at most ds_max); the flash is the hand-rolled
# model's, which compute_flash_black_oil_ad equals without vaporised oil.
# ---------------------------------------------------------------------------
_KNOWN3 = _KNOWN + ("grat",)


def control_form3(w: Well, qw: float, qo: float, qg: float):
    """(coefficients on qWs, qOs, qGs, bhp; target) of a three-phase well's
    control row (_control_equation, facility_model.py:2589-2658)."""
    if w.type not in _KNOWN3:
        raise NotImplementedError(f"well control '{w.type}' is not available on the device")
    total = (1.0, 1.0, 1.0, 0.0)
    if w.type == "bhp":
        qt = qw + qo + qg
        if (not getattr(w, "allow_sign_change", False) and w.sign != 0 and abs(qt) > 1e-8
                and np.sign(qt) != w.sign):
            return total, 0.0                     # flowing the wrong way: hold the total rate at zero
        return (0.0, 0.0, 0.0, 1.0), float(w.val)
    if float(w.val) == 0.0:
        return total, 0.0                         # a zero target closes on the total rate
    if w.type in _RATE_TYPES:
        return total, float(w.val)
    single = {"wrat": (1.0, 0.0, 0.0, 0.0), "orat": (0.0, 1.0, 0.0, 0.0), "grat": (0.0, 0.0, 1.0, 0.0),
              "lrat": (1.0, 1.0, 0.0, 0.0)}
    if w.type in single:
        return single[w.type], float(w.val)
    rho = getattr(w, "control_density", getattr(w, "ControlDensity", None))
    rho = np.ones(3) if rho is None else np.resize(np.asarray(rho, float).ravel(), 3)
    return (float(rho[0]), float(rho[1]), float(rho[2]), 0.0), float(w.val)      # resv


@dataclass
class GenericBOState(BOState):
    """Host live-oil state with the wells' perforation fluxes of the last
    converged step (MRST's wellSol.flux), one (nperf, 3) array per well."""
    well_flux: Optional[List[np.ndarray]] = None

    def copy(self):
        b = BOState.copy(self)
        return GenericBOState(b.pressure, b.s, b.rs, b.status, b.qWs, b.qOs, b.qGs, b.bhp,
                              None if self.well_flux is None else [f.copy() for f in self.well_flux])


@dataclass
class DeviceGenericBOState(DeviceBOState):
    well_flux: Optional[List[np.ndarray]] = None

    def copy(self):
        b = DeviceBOState.copy(self)
        return DeviceGenericBOState(b.p, b.s, b.rs, b.status, b.q, b.bhp,
                                    None if self.well_flux is None else [f.copy() for f in self.well_flux])

    def to_host(self) -> GenericBOState:
        h = DeviceBOState.to_host(self)
        return GenericBOState(h.pressure, h.s, h.rs, h.status, h.qWs, h.qOs, h.qGs, h.bhp,
                              None if self.well_flux is None else [f.copy() for f in self.well_flux])

    @staticmethod
    def from_host(st, device) -> "DeviceGenericBOState":
        d = DeviceBOState.from_host(st, device)
        flux = getattr(st, "well_flux", None)
        return DeviceGenericBOState(d.p, d.s, d.rs, d.status, d.q, d.bhp,
                                    None if flux is None else [np.array(f, float).reshape(-1, 3) for f in flux])


class GenericLiveOilModelAD(GenericBlackOilModelAD):
    """GenericBlackOilModelAD(g, rock, fluid, wells, disgas=True) with a
    deck fluid (init_deck_blackoil_fluid), on one GPU: made by
    GenericBlackOilModelAD itself when the fluid is a DeckBlackOilFluid.
    Wells may carry ``lims`` and ``status`` as in the oil-water form."""

    n_phases = 3

    def __init__(self, g: Grid, rock: Rock, fluid: DeckBlackOilFluid, wells: Sequence[Well],
                 disgas: bool = False, gravity: float = 0.0, ds_max: float = 0.2, dp_max_rel: float = 0.2,
                 tol_cnv: float = 1e-3, tol_mb: float = 1e-7, tol_well: float = 1e-6, bc=None, src=None,
                 vapoil: bool = False, aquifer=None, mass_basis: bool = False, operators=None,
                 linear_solver: Optional[CPRSolver] = None, device=None, allow_cross_flow: bool = True,
                 max_iterations: int = 25, drs_max: float = np.inf):
        if vapoil:
            raise NotImplementedError("vaporised oil is not available on the device")
        if not disgas:
            raise NotImplementedError("a deck fluid on the device is live oil: pass disgas=True "
                                      "(three-phase dead oil is not available)")
        if bc is not None or src is not None or aquifer is not None or operators is not None:
            raise NotImplementedError("boundary conditions, sources, aquifers and custom operators are not "
                                      "available on the device")
        if mass_basis:
            raise NotImplementedError("the device generic model assembles in surface volumes")
        if not allow_cross_flow:
            raise NotImplementedError("cross-flow is always allowed on the device (the reference's default)")
        self.wells = list(wells)
        for w in self.wells:
            control_form3(w, 0.0, 0.0, 0.0)      # rejects controls the device does not close
        legacy = [Well(cells=w.cells, type="bhp" if w.type == "bhp" else "rate", val=float(w.val), WI=w.WI,
                       dZ=w.dZ, compi=w.compi, sign=w.sign, name=w.name) for w in self.wells]
        self._dev = ThreePhaseBlackOilModelAD(g, rock, fluid, legacy, gravity=gravity, ds_max=ds_max,
                                              tol_cnv=tol_cnv, tol_mb=tol_mb, tol_well=tol_well,
                                              max_iterations=max_iterations, linear_solver=linear_solver,
                                              device=device)
        _lib.check(self._dev.lib.pmx_bo_set_update_rule(self._dev.h, 1, float(dp_max_rel), float(drs_max)),
                   "pmx_bo_set_update_rule")
        self.g, self.rock, self.fluid = g, rock, fluid
        self.gravity = float(gravity)
        self.ds_max, self.dp_max_rel, self.drs_max = ds_max, dp_max_rel, drs_max
        self.tol_cnv, self.tol_mb, self.tol_well = tol_cnv, tol_mb, tol_well
        self.max_iterations = max_iterations
        self.nc, self.nw = self._dev.nc, self._dev.nw
        self.device = self._dev.device
        self.disgas, self.vapoil = True, False
        self.within_limits = True
        self.switches = []
        self._rho_scale = np.ones(3)
        self._pmean = self._rsmean = 0.0
        self.last_cdp = None
        if self.nw:
            self._perf_cells = torch.as_tensor(np.concatenate([w.cells for w in self.wells]).astype(np.int64),
                                               device=self.device)

    def initial_state(self, p0, sw0, sg0, rs0=None) -> GenericBOState:
        b = self._dev.initial_state(p0, sw0, sg0, rs0)
        return GenericBOState(b.pressure, b.s, b.rs, b.status, b.qWs, b.qOs, b.qGs, b.bhp, None)

    def device_state(self, st) -> DeviceGenericBOState:
        """A device state from ours, a BOState-like host state, or the
        reference's mapping State (pressure, sW, sG, rs, status, qWs, qOs,
        qGs, bhp and optionally well_flux)."""
        if isinstance(st, DeviceGenericBOState):
            return st
        if isinstance(st, DeviceBOState):
            return DeviceGenericBOState(st.p, st.s, st.rs, st.status, st.q, st.bhp, None)
        if hasattr(st, "keys"):
            sw, sg = np.asarray(st["sW"], float), np.asarray(st["sG"], float)
            so = np.asarray(st["sO"], float) if "sO" in st else 1.0 - sw - sg
            st = GenericBOState(np.asarray(st["pressure"], float), np.column_stack([sw, so, sg]),
                                np.asarray(st["rs"], float), np.asarray(st["status"], np.int64),
                                np.asarray(st["qWs"], float), np.asarray(st["qOs"], float),
                                np.asarray(st["qGs"], float), np.asarray(st["bhp"], float), st.get("well_flux"))
        return DeviceGenericBOState.from_host(st, self.device)

    # -- per ministep: accumulation, rhoScale and the column drops at state0 ---
    def begin_ministep(self, state0: DeviceGenericBOState):
        dev = self._dev
        dev.begin_step(state0)
        P = _lib.ptr
        out = torch.empty(6 * self.nc, dtype=torch.float64, device=self.device)
        _lib.check(dev.lib.pmx_bo_cell_densities(dev.h, P(state0.p), P(state0.s), P(state0.rs), P(state0.status),
                                                 P(out)), "pmx_bo_cell_densities")
        props = out.view(self.nc, 6)
        mean = props[:, :3].mean(dim=0).cpu().numpy()
        rhoS = np.asarray(self.fluid.rhoS, float)
        self._rho_scale = np.array([rhoS[i] / mean[i] if np.isfinite(mean[i]) and mean[i] > 0 else 1.0
                                    for i in range(3)])
        if self.gravity and self.nw:
            perf = props[self._perf_cells].cpu().numpy()
            out_w = []
            for k, w in enumerate(self.wells):
                a, b = dev.perf_ptr[k], dev.perf_ptr[k + 1]
                flux = None if state0.well_flux is None else state0.well_flux[k]
                out_w.append(column_drops(w.dZ, np.asarray(w.WI, float), perf[a:b, :3], perf[a:b, 3:], w.compi,
                                          w.sign, flux, self.gravity))
            cdp = np.ascontiguousarray(np.concatenate(out_w))
            _lib.check(dev.lib.pmx_bo_set_connection_drops(dev.h, P(cdp)), "pmx_bo_set_connection_drops")
            self.last_cdp = out_w

    # -- one Newton iteration: controls, assembly, convergence -----------------
    def _set_controls(self, st: DeviceGenericBOState):
        q = st.q.cpu().numpy().reshape(-1, 3) if self.nw else np.zeros((0, 3))
        coef, val = np.zeros((max(self.nw, 1), 4)), np.zeros(max(self.nw, 1))
        for k, w in enumerate(self.wells):
            c, v = control_form3(w, float(q[k, 0]), float(q[k, 1]), float(q[k, 2]))
            coef[k], val[k] = c, v
        status = np.array([1 if getattr(w, "status", True) else 0 for w in self.wells] or [1], dtype=np.int32)
        coef = np.ascontiguousarray(coef.ravel())
        rs = np.ascontiguousarray(self._rho_scale, dtype=np.float64)
        _lib.check(self._dev.lib.pmx_bo_set_controls(self._dev.h, _lib.ptr(coef), _lib.ptr(val), _lib.ptr(status),
                                                     _lib.ptr(rs)), "pmx_bo_set_controls")

    def linearize(self, state0: DeviceGenericBOState, st: DeviceGenericBOState, dt: float):
        """get_equations + get_convergence: (converged, {CNV, MB, well, control})."""
        self.within_limits = True
        self._set_controls(st)
        dev = self._dev
        dev.assemble_device(st, dt)
        out, wr = dev.convergence_values(st)
        nc = self.nc
        b_avg = out[6:9] / nc
        cnv = max(b_avg[a] * dt * out[a] for a in range(3))
        mb = max(b_avg[a] * dt * abs(out[3 + a]) / out[11] for a in range(3))
        well_res, ctrl_res = self._well_residuals(wr)
        self._pmean, self._rsmean = out[9] / nc, out[10] / nc
        ok = (cnv <= self.tol_cnv and mb <= self.tol_mb and well_res <= self.tol_well and ctrl_res <= self.tol_well
              and self.within_limits)
        return bool(ok), {"CNV": float(cnv), "MB": float(mb), "well": float(well_res), "control": float(ctrl_res)}

    def _well_residuals(self, wr: np.ndarray):
        """convergence.py:72-112 on the three rate rows and the control row
        of every well."""
        if not self.nw:
            return 0.0, 0.0
        wr = np.asarray(wr, float).reshape(self.nw, 4)
        rates = [abs(w.val) for w in self.wells if w.type != "bhp" and abs(w.val) > 1e-12]
        q_ref = max(rates) if rates else 1.0
        well = np.abs(wr[:, :3]).max() / q_ref
        scale = np.array([1e5 if w.type == "bhp" else (abs(w.val) if abs(w.val) > 1e-12 else max(q_ref, 1.0))
                          for w in self.wells])
        return float(well), float(np.max(np.abs(wr[:, 3]) / scale))

    # -- Newton update and control switching ----------------------------------
    def _switch_controls(self, st: DeviceGenericBOState):
        if not any(getattr(w, "lims", None) for w in self.wells):
            return
        bhp = st.bhp.cpu().numpy()
        q = st.q.cpu().numpy().reshape(-1, 3)
        for k, w in enumerate(self.wells):
            if not getattr(w, "status", True) or not getattr(w, "lims", None):
                continue
            mode = limit_broken(w, float(bhp[k]), q[k])
            if mode is None:
                continue
            self.switches.append((k, w.type, mode))
            w.type, w.val = mode, float(w.lims[mode])
            self.within_limits = False

    def newton_update(self, st: DeviceGenericBOState, scale: float = 1.0) -> bool:
        self._dev.stats["newton"] += 1
        ok = self._dev.solve_and_update(st, self._pmean, self._rsmean, scale)
        if ok:
            self._switch_controls(st)
        return ok

    def apply_increment(self, st: DeviceGenericBOState, scale: float = 1.0):
        self._dev.apply_increment(st, scale)
        self._switch_controls(st)

    def after_convergence(self, st: DeviceGenericBOState, state0, dt: float) -> DeviceGenericBOState:
        """The converged perforation fluxes go with the state, for the next
        step's wellbore column."""
        if self.nw:
            dev = self._dev
            rates = np.zeros(3 * int(dev.perf_ptr[-1]))
            _lib.check(dev.lib.pmx_bo_perf_volume_rates(dev.h, _lib.ptr(rates)), "pmx_bo_perf_volume_rates")
            st.well_flux = [rates.reshape(-1, 3)[dev.perf_ptr[k]:dev.perf_ptr[k + 1]].copy()
                            for k in range(self.nw)]
        return st
