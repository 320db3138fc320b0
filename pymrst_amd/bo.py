"""Three-phase live-oil black-oil model on the MI355X: the reference's
ThreePhaseBlackOilModelAD call surface over the pmx HIP library.

Reference: pymrst/modules/ad_blackoil/models/three_phase_black_oil_model.py
(BOState :74-91, __init__ :116-160, assemble :314-381, _well :383-436,
update_state :439-470, _flash :472-509, _conv_scaling/_converged :512-545,
step :547-571) and the deck fluid of
pymrst/modules/ad_props/props/init_deck_adi_fluid.py:60-194.

Residual, Jacobian (3x3 blocks plus the well tail), CPR-FGMRES solve and the
update with the variable switch and flash run on the GPU; the host keeps the
Newton control flow (convergence decisions from device reductions).
"""
import ctypes
import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from .grid import BARSA, NORM_GRAVITY, Grid, Rock, Well, cartesian_lines, compute_trans, pore_volume, \
    tpfa_connections
from .ow import CPRSolver


@dataclass
class DeckBlackOilFluid:
    """The deck keywords init_deck_blackoil_fluid reads (SI units):
    ``rhoS`` (water, oil, gas surface densities), ``pvtw`` (p_ref, b_ref, c,
    mu_ref, viscosibility), ``rock_cp`` (p_ref, c), ``pvdg`` rows (p, Bg,
    mu_g), ``pvto`` as a list of (Rs, rows (p, Bo, mu_o)) with the saturated
    row first, ``swof`` and ``sgof`` rows (s, kr, kro, pc).  The tables are
    built on the device as the reference builds them: PVDG with a leading
    row at p = 0, rs_sat through (0, 0), extend_tab on SWOF/SGOF, the
    saturation-weighted three-phase oil relperm (init_deck_adi_fluid.py:52-57)."""
    rhoS: np.ndarray
    pvtw: dict
    rock_cp: dict
    pvdg: np.ndarray
    pvto: list
    swof: np.ndarray
    sgof: np.ndarray

    @property
    def swcon(self) -> float:
        return float(self.swof[0, 0])

    def abi_arrays(self):
        w, r = self.pvtw, self.rock_cp
        scalars = np.array([self.rhoS[0], self.rhoS[1], self.rhoS[2], w["p_ref"], w["b_ref"], w["c"], w["mu_ref"],
                            r["p_ref"], r["c"]], dtype=np.float64)
        keys = np.array([k for k, _ in self.pvto], dtype=np.float64)
        rows = [np.asarray(b, np.float64).reshape(-1, 3) for _, b in self.pvto]
        bptr = np.concatenate([[0], np.cumsum([b.shape[0] for b in rows])]).astype(np.int32)
        branch = np.ascontiguousarray(np.vstack(rows))
        pvdg = np.ascontiguousarray(self.pvdg, dtype=np.float64)
        swof = np.ascontiguousarray(self._four(self.swof))
        sgof = np.ascontiguousarray(self._four(self.sgof))
        return scalars, pvdg, keys, bptr, branch, swof, sgof

    @staticmethod
    def _four(t):
        t = np.asarray(t, np.float64)
        if t.shape[1] == 4:
            return t
        out = np.zeros((t.shape[0], 4))
        out[:, :t.shape[1]] = t
        return out

    # host evaluation of the saturation curve (initial states)
    def rs_sat(self, p):
        p_t = np.concatenate([[0.0], [np.asarray(b, float)[0, 0] for _, b in self.pvto]])
        rs_t = np.concatenate([[0.0], [k for k, _ in self.pvto]])
        p = np.asarray(p, float)
        idx = np.clip(np.searchsorted(p_t, p, side="right") - 1, 0, p_t.size - 2)
        t = (p - p_t[idx]) / (p_t[idx + 1] - p_t[idx])
        return rs_t[idx] * (1 - t) + rs_t[idx + 1] * t


def init_deck_blackoil_fluid(deck: dict, rock=None) -> DeckBlackOilFluid:
    """init_deck_blackoil_fluid (init_deck_adi_fluid.py:60-194) for the
    dissolved-gas decks it supports with PVDG gas (PVTG wet gas, PVDO dead
    oil and J-function scaling are not part of the device path)."""
    if deck["pvtw"].get("viscosibility", 0.0) != 0.0:
        raise ValueError("water viscosibility not supported")
    for k in ("pvdg", "pvto", "swof", "sgof", "rhoS", "rock_cp"):
        if k not in deck:
            raise NotImplementedError(f"deck keyword '{k}' is required by the device black-oil fluid")
    if deck.get("jfunc") is not None or deck.get("JFUNC") is not None:
        raise NotImplementedError("J-function scaling is not supported by the device black-oil fluid")
    return DeckBlackOilFluid(rhoS=np.asarray(deck["rhoS"], float), pvtw=dict(deck["pvtw"]),
                             rock_cp=dict(deck["rock_cp"]), pvdg=np.asarray(deck["pvdg"], float),
                             pvto=[(float(k), np.asarray(v, float)) for k, v in deck["pvto"]],
                             swof=np.asarray(deck["swof"], float), sgof=np.asarray(deck["sgof"], float))


@dataclass
class BOState:
    """Host state, as the reference's BOState (:74-91)."""
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


@dataclass
class DeviceBOState:
    """The same state in device tensors: p (nc), s (nc, 3), rs (nc),
    status (nc, int32), q (nw, 3) = (qWs, qOs, qGs), bhp (nw)."""
    p: torch.Tensor
    s: torch.Tensor
    rs: torch.Tensor
    status: torch.Tensor
    q: torch.Tensor
    bhp: torch.Tensor

    def copy(self):
        return DeviceBOState(*(t.clone() for t in (self.p, self.s, self.rs, self.status, self.q, self.bhp)))

    def to_host(self) -> BOState:
        q = self.q.cpu().numpy()
        return BOState(self.p.cpu().numpy(), self.s.cpu().numpy(), self.rs.cpu().numpy(),
                       self.status.cpu().numpy().astype(np.int64), q[:, 0].copy(), q[:, 1].copy(), q[:, 2].copy(),
                       self.bhp.cpu().numpy())

    @staticmethod
    def from_host(st: BOState, device) -> "DeviceBOState":
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
        q = np.column_stack([st.qWs, st.qOs, st.qGs]) if np.size(st.qWs) else np.zeros((0, 3))
        return DeviceBOState(f(st.pressure), f(np.asarray(st.s).reshape(-1, 3)), f(st.rs),
                             torch.as_tensor(np.ascontiguousarray(st.status, dtype=np.int32), device=device),
                             f(q), f(st.bhp))


class ThreePhaseBlackOilModelAD:
    """Drop-in for the reference's ThreePhaseBlackOilModelAD on one GPU.

    Differences from the reference, all deliberate:
    * the Newton increment comes from the device CPR-FGMRES (``linear_solver``
      is a :class:`CPRSolver`; the reference solves directly); a linear solve
      that does not converge makes ``step`` report failure;
    * the fluid is the deck fluid of :func:`init_deck_blackoil_fluid`.
    """

    n_phases = 3

    def __init__(self, g: Grid, rock: Rock, fluid: DeckBlackOilFluid, wells: Sequence[Well],
                 gravity: float = NORM_GRAVITY, ds_max: float = 0.2, tol_cnv: float = 1e-3, tol_mb: float = 1e-7,
                 tol_well: float = 1e-6, max_iterations: int = 25, linear_solver: Optional[CPRSolver] = None,
                 device: Optional[torch.device] = None, T: Optional[np.ndarray] = None,
                 N: Optional[np.ndarray] = None):
        if not torch.cuda.is_available():
            raise _lib.PmxError("ThreePhaseBlackOilModelAD needs a GPU (no CPU fallback)")
        self.lib = _lib.load()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.g, self.rock, self.fluid = g, rock, fluid
        self.wells = list(wells)
        for w in self.wells:
            if w.type not in ("bhp", "rate"):
                raise ValueError(f"unsupported well control '{w.type}'")
        own_connections = N is None or T is None
        if own_connections:
            N, T, _, _ = tpfa_connections(g, compute_trans(g, rock))
        self.N = np.ascontiguousarray(N, dtype=np.int32)
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        self.pv_ref = pore_volume(g, rock)
        z = g.cell_centroids[:, 2]
        self.gravity = float(gravity)
        self.gdz = gravity * (z[self.N[:, 1]] - z[self.N[:, 0]]) if gravity else np.zeros(self.N.shape[0])
        self.ds_max = ds_max
        self.tol_cnv, self.tol_mb, self.tol_well = tol_cnv, tol_mb, tol_well
        self.max_iterations = max_iterations
        self.linear_solver = linear_solver if linear_solver is not None else CPRSolver()
        self.nc = g.num_cells
        self.nw = len(self.wells)
        self.stats = {"newton": 0, "linear": 0, "assemble_s": 0.0, "solve_s": 0.0}
        P = _lib.ptr
        with torch.cuda.device(self.device):
            self.stream = torch.cuda.current_stream(self.device)
            sc, pvdg, keys, bptr, branch, swof, sgof = fluid.abi_arrays()
            h = ctypes.c_void_p()
            _lib.check(self.lib.pmx_bo_create(self.device.index or 0, ctypes.c_void_p(self.stream.cuda_stream),
                                              P(sc), pvdg.shape[0], P(pvdg), keys.size, P(keys), P(bptr),
                                              P(branch), swof.shape[0], P(swof), sgof.shape[0], P(sgof),
                                              ctypes.byref(h)), "pmx_bo_create")
            self.h = h
            pv = np.ascontiguousarray(self.pv_ref, dtype=np.float64)
            gdz = np.ascontiguousarray(self.gdz, dtype=np.float64)
            _lib.check(self.lib.pmx_bo_set_topology(h, self.nc, self.N.shape[0], P(self.N), P(self.T), P(gdz),
                                                    P(pv)), "pmx_bo_set_topology")
            self._set_wells()
            if own_connections and getattr(g, "cartDims", None) is not None:
                _, ptr, cells = cartesian_lines(g, self.N, self.T)
                _lib.check(self.lib.pmx_bo_set_lines(h, ptr.size - 1, P(ptr), P(cells)), "pmx_bo_set_lines")
        self._dx = torch.zeros(3 * self.nc + 4 * self.nw, dtype=torch.float64, device=self.device)

    def _set_wells(self):
        nw = self.nw
        counts = [w.cells.size for w in self.wells]
        self.perf_ptr = np.ascontiguousarray(np.concatenate([[0], np.cumsum(counts)]), dtype=np.int32)
        cells = (np.ascontiguousarray(np.concatenate([w.cells for w in self.wells]), dtype=np.int32)
                 if nw else np.zeros(0, np.int32))
        wi = np.ascontiguousarray(np.concatenate([w.WI for w in self.wells]), dtype=np.float64) if nw else np.zeros(0)
        types = np.array([0 if w.type == "bhp" else 1 for w in self.wells], dtype=np.int32)
        vals = np.array([w.val for w in self.wells], dtype=np.float64)
        compi = np.zeros((nw, 3))
        for k, w in enumerate(self.wells):
            c = np.asarray(w.compi, float).ravel()
            compi[k, :min(3, c.size)] = c[:3]
        compi = np.ascontiguousarray(compi.reshape(-1))
        P = _lib.ptr
        _lib.check(self.lib.pmx_bo_set_wells(self.h, nw, P(self.perf_ptr), P(cells), P(wi), P(types), P(vals),
                                             P(compi)), "pmx_bo_set_wells")
        self._q_ref = max((abs(w.val) for w in self.wells), default=1.0)
        self._ctrl_scale = np.array([1e5 if w.type == "bhp" else abs(w.val) for w in self.wells])

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and getattr(self, "lib", None) is not None:
            try:
                self.lib.pmx_bo_destroy(h)
            except Exception:
                pass
            self.h = None

    # -- reference attribute contract -----------------------------------------
    @property
    def G(self):
        return self.g

    @property
    def nnzb(self) -> int:
        v = ctypes.c_int32()
        _lib.check(self.lib.pmx_bo_nnzb(self.h, ctypes.byref(v)), "pmx_bo_nnzb")
        return int(v.value)

    def reset(self):
        self.stats = {"newton": 0, "linear": 0, "assemble_s": 0.0, "solve_s": 0.0}

    def profile(self, enable: bool):
        _lib.check(self.lib.pmx_bo_profile(self.h, int(enable)), "pmx_bo_profile")

    def launch_shapes(self) -> dict:
        return _lib.read_launch_shapes(self.lib.pmx_bo_launch_shapes, self.h)

    def profile_read(self) -> dict:
        out = np.zeros(4)
        _lib.check(self.lib.pmx_bo_profile_read(self.h, _lib.ptr(out)), "pmx_bo_profile_read")
        return {"spmv_count": int(out[0]), "spmv_ms": float(out[1]), "prec_count": int(out[2]),
                "prec_ms": float(out[3])}

    def initial_state(self, p0, sw0, sg0, rs0=None) -> BOState:
        """Cells with free gas start saturated (status 3, rs on the
        saturation curve); gas-free cells start undersaturated at rs0
        (default: saturated Rs at p0, capped by it)."""
        nc, nw = self.nc, self.nw
        p = np.full(nc, float(p0)) if np.size(p0) == 1 else np.asarray(p0, float).copy()
        sw = np.broadcast_to(np.asarray(sw0, float), (nc,)).copy()
        sg = np.broadcast_to(np.asarray(sg0, float), (nc,)).copy()
        rsat = self.fluid.rs_sat(p)
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

    def device_state(self, st) -> DeviceBOState:
        return st if isinstance(st, DeviceBOState) else DeviceBOState.from_host(st, self.device)

    # -- device primitives -----------------------------------------------------
    def begin_step(self, st0: DeviceBOState):
        P = _lib.ptr
        _lib.check(self.lib.pmx_bo_begin_step(self.h, P(st0.p), P(st0.s), P(st0.rs), P(st0.status)),
                   "pmx_bo_begin_step")

    def assemble_device(self, st: DeviceBOState, dt: float):
        P = _lib.ptr
        _lib.check(self.lib.pmx_bo_assemble(self.h, P(st.p), P(st.s), P(st.rs), P(st.status), P(st.q), P(st.bhp),
                                            float(dt)), "pmx_bo_assemble")

    def residual_device(self) -> torch.Tensor:
        r = torch.empty(3 * self.nc + 4 * self.nw, dtype=torch.float64, device=self.device)
        _lib.check(self.lib.pmx_bo_residual(self.h, _lib.ptr(r)), "pmx_bo_residual")
        return r

    def matvec(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty_like(x)
        _lib.check(self.lib.pmx_bo_matvec(self.h, _lib.ptr(x), _lib.ptr(y)), "pmx_bo_matvec")
        return y

    def convergence_values(self, st: DeviceBOState):
        out = np.zeros(12)
        wells = np.zeros(max(4 * self.nw, 1))
        _lib.check(self.lib.pmx_bo_convergence(self.h, _lib.ptr(st.p), _lib.ptr(st.rs), _lib.ptr(st.status),
                                               _lib.ptr(out), _lib.ptr(wells)), "pmx_bo_convergence")
        return out, wells[:4 * self.nw]

    def converged(self, dt: float, st: DeviceBOState):
        """_converged (:522-545) from device reductions: CNV and MB per phase
        with B_avg = mean(1/b_a), well equations against tol_well * max|target|,
        controls against tol_well * (1e5 for bhp, |target| for rate).
        Returns (ok, pmean, rsmean)."""
        out, wr = self.convergence_values(st)
        nc = self.nc
        b_avg = out[6:9] / nc
        cnv = max(b_avg[a] * dt * out[a] for a in range(3))
        mb = max(b_avg[a] * dt * abs(out[3 + a]) / out[11] for a in range(3))
        ok = bool(cnv <= self.tol_cnv and mb <= self.tol_mb)
        if self.nw:
            wr = wr.reshape(self.nw, 4)
            well_ok = bool(np.all(np.abs(wr[:, :3]) <= self.tol_well * self._q_ref))
            ctrl_ok = bool(np.all(np.abs(wr[:, 3]) <= self.tol_well * self._ctrl_scale))
            ok = ok and well_ok and ctrl_ok
        self.last_cnv_mb = (cnv, mb)
        return ok, out[9] / nc, out[10] / nc

    def solve_and_update(self, st: DeviceBOState, pmean: float, rsmean: float, scale: float = 1.0) -> bool:
        info = _lib.SolveInfo()
        opts = self.linear_solver.options()
        t0 = time.perf_counter()
        rc = self.lib.pmx_bo_solve(self.h, ctypes.byref(opts), float(pmean), float(rsmean), _lib.ptr(self._dx),
                                   ctypes.byref(info))
        self.stats["solve_s"] += time.perf_counter() - t0
        _lib.check(rc, "pmx_bo_solve")
        self.linear_solver.last_iterations = int(info.iterations)
        self.linear_solver.history.append((int(info.iterations), float(info.rel_residual), float(info.setup_ms),
                                           float(info.solve_ms)))
        self.stats["linear"] += int(info.iterations)
        if rc != 0:
            return False
        self.apply_increment(st, scale)
        return True

    def apply_increment(self, st: DeviceBOState, scale: float = 1.0):
        """update_state(st, scale * (the last Newton increment)), with the
        variable switch and the flash."""
        P = _lib.ptr
        dx = self._dx if scale == 1.0 else self._dx * float(scale)
        _lib.check(self.lib.pmx_bo_update(self.h, P(dx), float(self.ds_max), P(st.p), P(st.s), P(st.rs),
                                          P(st.status), P(st.q), P(st.bhp)), "pmx_bo_update")

    # -- ministep protocol of pymrst_amd.nonlinear.NonLinearSolver --------------
    def begin_ministep(self, state0: DeviceBOState):
        self.begin_step(state0)

    def linearize(self, state0: DeviceBOState, st: DeviceBOState, dt: float):
        """Assemble at st and test convergence: (converged, {CNV, MB})."""
        self.assemble_device(st, dt)
        ok, self._ms_pmean, self._ms_rsmean = self.converged(dt, st)
        cnv, mb = self.last_cnv_mb
        return ok, {"CNV": float(cnv), "MB": float(mb)}

    def newton_update(self, st: DeviceBOState, scale: float = 1.0) -> bool:
        self.stats["newton"] += 1
        return self.solve_and_update(st, self._ms_pmean, self._ms_rsmean, scale)

    @property
    def last_linear_solve(self):
        h = self.linear_solver.history[-1]
        return int(h[0]), float(h[1])

    # -- Newton (step :547-571) --------------------------------------------------
    def step_device(self, state: DeviceBOState, dt: float):
        st = state.copy()
        self.begin_step(state)
        for it in range(self.max_iterations):
            t0 = time.perf_counter()
            self.assemble_device(st, dt)
            ok, pmean, rsmean = self.converged(dt, st)
            self.stats["assemble_s"] += time.perf_counter() - t0
            if ok:
                return st, True, it
            self.stats["newton"] += 1
            if not self.solve_and_update(st, pmean, rsmean):
                return st, False, it + 1
        self.assemble_device(st, dt)
        ok, _, _ = self.converged(dt, st)
        return st, ok, self.max_iterations

    def step(self, state, dt: float):
        if isinstance(state, DeviceBOState):
            return self.step_device(state, dt)
        st, ok, its = self.step_device(DeviceBOState.from_host(state, self.device), dt)
        return st.to_host(), ok, its
