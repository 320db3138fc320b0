// Fully-implicit three-phase live-oil black-oil residual and Jacobian on the
// device.
//
// Equations (pymrst/modules/ad_blackoil/models/three_phase_black_oil_model.py
// assemble :314-381, _well :383-436, update_state :439-470, _flash :472-509,
// _conv_scaling/_converged :512-545) with the flux functions of
// ad_blackoil/utils/get_flux_and_props_{water,oil,gas}_bo.py and the deck
// fluid of ad_props/props/init_deck_adi_fluid.py:52-194:
//   R_w = (pv b_w s_w - [pv b_w s_w]^0)/dt + div(upw(b_w) v_w) - q_w
//   R_o = (pv b_o s_o - [pv b_o s_o]^0)/dt + div(upw(b_o) v_o) - q_o
//   R_g = (pv (b_g s_g + rs b_o s_o) - [..]^0)/dt
//         + div(upw(b_g) v_g + upw(rs) upw(b_o) v_o) - q_g
//   v_a = -(kr_a/mu_a)_up T (dp - avg(rho_a) gdz),  upwind flag dp <= 0 -> N0.
// Cell unknowns (p, sW, x) with x = rs in undersaturated cells (status 1)
// and x = sG in saturated cells (status 3); well unknowns (qWs, qOs, qGs,
// bhp).  Derivatives come from forward-mode dual numbers over one cell's
// three unknowns (and the well's bhp), evaluated in registers.
//
// Layout: cell-major interleaved unknowns and residual rows (water, oil,
// gas); 3x3 row-major blocks [d eq / d (p, sW, x)]; per-cell properties with
// their derivatives are written once per assembly (28 doubles per cell) and
// gathered by the face loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "bo.h"
#include "ow_props.h"

namespace pmx {

// ---------------------------------------------------------------------------
// dual numbers
// ---------------------------------------------------------------------------
template <int N>
struct Dl {
  double v;
  double d[N];
};

template <int N>
__host__ __device__ __forceinline__ Dl<N> dc(double v) {
  Dl<N> r;
  r.v = v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> dvar(double v, int i) {
  Dl<N> r = dc<N>(v);
  r.d[i] = 1.0;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator+(const Dl<N>& a, const Dl<N>& b) {
  Dl<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator-(const Dl<N>& a, const Dl<N>& b) {
  Dl<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator-(double a, const Dl<N>& b) {
  Dl<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator*(const Dl<N>& a, const Dl<N>& b) {
  Dl<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator*(double a, const Dl<N>& b) {
  Dl<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator/(const Dl<N>& a, const Dl<N>& b) {
  Dl<N> r;
  const double inv = 1.0 / b.v;
  r.v = a.v / b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator/(const Dl<N>& a, double b) {
  Dl<N> r;
  r.v = a.v / b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / b;
  return r;
}
template <int N>
__host__ __device__ __forceinline__ Dl<N> operator/(double a, const Dl<N>& b) {
  Dl<N> r;
  r.v = a / b.v;
  const double s = -r.v / b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = s * b.d[i];
  return r;
}
// Chain rule through a scalar function with value f and slope df.
template <int N>
__host__ __device__ __forceinline__ Dl<N> chain(double f, double df, const Dl<N>& x) {
  Dl<N> r;
  r.v = f;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = df * x.d[i];
  return r;
}
// Embed a 3-variable dual at offset O of a wider one.
template <int N, int O>
__host__ __device__ __forceinline__ Dl<N> lift(const Dl<3>& a) {
  Dl<N> r = dc<N>(a.v);
#pragma unroll
  for (int i = 0; i < 3; ++i) r.d[O + i] = a.d[i];
  return r;
}

// ---------------------------------------------------------------------------
// tables (interp_table.py:17-59 and interp_pvt.py:41-70, 175-230)
// ---------------------------------------------------------------------------
// numpy.searchsorted(x, v, 'right') - 1, clipped to [0, n-2]
__host__ __device__ __forceinline__ int tbin(const double* x, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (x[m] <= v)
      lo = m + 1;
    else
      hi = m;
  }
  int b = lo - 1;
  b = b < 0 ? 0 : b;
  return b > n - 2 ? n - 2 : b;
}
// MATLAB interp1 weights form with linear extrapolation: f0 (1-t) + f1 t.
__host__ __device__ __forceinline__ void interp_w(const double* x, const double* f, int n, double v, double& y,
                                                  double& dy) {
  const int b = tbin(x, n, v);
  const double x0 = x[b], x1 = x[b + 1], f0 = f[b], f1 = f[b + 1];
  const double t = (v - x0) / (x1 - x0);
  y = f0 * (1.0 - t) + f1 * t;
  dy = (f1 - f0) / (x1 - x0);
}
// Slope form of the PVT tables: f0 + slope (v - x0).
__host__ __device__ __forceinline__ void interp_s(const double* x, const double* f, int n, double v, double& y,
                                                  double& dy) {
  const int b = tbin(x, n, v);
  const double x0 = x[b], x1 = x[b + 1], f0 = f[b], f1 = f[b + 1];
  const double slope = (f1 - f0) / (x1 - x0);
  y = f0 + slope * (v - x0);
  dy = slope;
}

struct PVTO2 {
  double b, b_p, b_r, mu, mu_p, mu_r;   // 1/Bo, mu_o and their d/dp, d/drs
};

// pvt_interp 'parallel' for 1/Bo and mu_o together (same bins).
__host__ __device__ __forceinline__ PVTO2 pvto(const BOFluidView& F, double p, double rs, bool sat) {
  const double* t = F.t;
  PVTO2 o;
  if (sat && !(p < t[F.sat_p])) {
    interp_s(t + F.sat_p, t + F.sat_b, F.nkey, p, o.b, o.b_p);
    interp_s(t + F.sat_p, t + F.sat_mu, F.nkey, p, o.mu, o.mu_p);
    o.b_r = o.mu_r = 0.0;
    return o;
  }
  const int b = tbin(t + F.key, F.nkey, rs);
  const double k0 = t[F.key + b], k1 = t[F.key + b + 1];
  const double w = (rs - k0) / (k1 - k0), dw = 1.0 / (k1 - k0);
  double ps, dps;
  interp_s(t + F.key, t + F.sat_p, F.nkey, rs, ps, dps);
  const double dp = p - ps;
  // branch b+1 weighted by w ("left"), branch b by 1-w ("right")
  const int l0 = F.bptr[b + 1], ln = F.bptr[b + 2] - l0;
  const int r0 = F.bptr[b], rn = F.bptr[b + 1] - r0;
  double fl, dl, fr, dr;
  interp_s(t + F.br_dp + l0, t + F.br_b + l0, ln, dp, fl, dl);
  interp_s(t + F.br_dp + r0, t + F.br_b + r0, rn, dp, fr, dr);
  o.b = fl * w + fr * (1.0 - w);
  o.b_p = dl * w + dr * (1.0 - w);
  o.b_r = (fl - fr) * dw - o.b_p * dps;
  interp_s(t + F.br_dp + l0, t + F.br_mu + l0, ln, dp, fl, dl);
  interp_s(t + F.br_dp + r0, t + F.br_mu + r0, rn, dp, fr, dr);
  o.mu = fl * w + fr * (1.0 - w);
  o.mu_p = dl * w + dr * (1.0 - w);
  o.mu_r = (fl - fr) * dw - o.mu_p * dps;
  return o;
}

__host__ __device__ __forceinline__ double quad(double c, double dp) {
  const double x = c * dp;
  return 1.0 + x + 0.5 * x * x;
}
__host__ __device__ __forceinline__ double rs_sat(const BOFluidView& F, double p, double& d) {
  double y;
  interp_w(F.t + F.rs_p, F.t + F.rs_k, F.n_rs, p, y, d);
  return y;
}
__host__ __device__ __forceinline__ double b_gas(const BOFluidView& F, double p, double& d) {
  double y;
  interp_w(F.t + F.g_p, F.t + F.g_b, F.n_g, p, y, d);
  return y;
}
__host__ __device__ __forceinline__ double b_wat(const BOFluidView& F, double p, double& d) {
  const double x = F.cw * (p - F.pw_ref);
  d = F.cw * (1.0 + x) / F.bw_ref;
  return (1.0 + x + 0.5 * x * x) / F.bw_ref;
}

// Cell properties with derivatives in the cell's (p, sW, x).
struct BOCell {
  Dl<3> bw, mw, bo, mo, rs, bg, mg;
};
constexpr int kBOProps = 28;

__device__ __forceinline__ BOCell bo_cell(const BOFluidView& F, double p, double sw, double sg_state,
                                          double rs_state, int status) {
  const double* t = F.t;
  const bool st1 = status == 1, st3 = status == 3;
  const Dl<3> P = dvar<3>(p, 0), SW = dvar<3>(sw, 1);
  const Dl<3> X = dvar<3>(st1 ? rs_state : sg_state, 2);
  BOCell c;
  double y, dy;
  if (st1) {
    c.rs = X;
  } else {
    y = rs_sat(F, p, dy);
    c.rs = chain(y, dy, P);
  }
  const Dl<3> SG = st3 ? X : dc<3>(0.0);
  const Dl<3> SO = 1.0 - SW - SG;
  y = b_wat(F, p, dy);
  c.bw = chain(y, dy, P);
  // water relperm / constant viscosity
  interp_w(t + F.w_s, t + F.w_kr, F.n_w, sw, y, dy);
  c.mw = chain(y, dy, SW) / F.muw;
  // live oil
  const PVTO2 o = pvto(F, p, c.rs.v, st3);
  c.bo = chain(o.b, o.b_p, P) + chain(0.0, o.b_r, c.rs);
  const Dl<3> muo = chain(o.mu, o.mu_p, P) + chain(0.0, o.mu_r, c.rs);
  // gas
  y = b_gas(F, p, dy);
  c.bg = chain(y, dy, P);
  interp_w(t + F.g_p, t + F.g_mu, F.n_g, p, y, dy);
  const Dl<3> mug = chain(y, dy, P);
  interp_w(t + F.gg_s, t + F.gg_kr, F.n_gg, SG.v, y, dy);
  c.mg = chain(y, dy, SG) / mug;
  // saturation-weighted oil relperm (init_deck_adi_fluid.py:52-57)
  const double swc = fmin(F.swcon, sw - 1e-5);
  const Dl<3> dd = SG + SW - dc<3>(swc);
  const Dl<3> ww = (SW - dc<3>(swc)) / dd;
  interp_w(t + F.og_s, t + F.og_kr, F.n_og, SO.v, y, dy);
  const Dl<3> kog = chain(y, dy, SO);
  interp_w(t + F.ow_s, t + F.ow_kr, F.n_ow, SO.v, y, dy);
  const Dl<3> kow = chain(y, dy, SO);
  c.mo = ((1.0 - ww) * kog + ww * kow) / muo;
  return c;
}

__device__ __forceinline__ void store_cell(double* __restrict__ dst, const BOCell& c) {
  const Dl<3>* src[7] = {&c.bw, &c.mw, &c.bo, &c.mo, &c.rs, &c.bg, &c.mg};
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    dst[4 * q] = src[q]->v;
#pragma unroll
    for (int i = 0; i < 3; ++i) dst[4 * q + 1 + i] = src[q]->d[i];
  }
}
__device__ __forceinline__ BOCell load_cell(const double* __restrict__ src) {
  double a[kBOProps];
  const double2* s2 = reinterpret_cast<const double2*>(src);
#pragma unroll
  for (int q = 0; q < kBOProps / 2; ++q) {
    const double2 u = s2[q];
    a[2 * q] = u.x;
    a[2 * q + 1] = u.y;
  }
  BOCell c;
  Dl<3>* dst[7] = {&c.bw, &c.mw, &c.bo, &c.mo, &c.rs, &c.bg, &c.mg};
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    dst[q]->v = a[4 * q];
#pragma unroll
    for (int i = 0; i < 3; ++i) dst[q]->d[i] = a[4 * q + 1 + i];
  }
  return c;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bo_props(int nc, const double* __restrict__ p, const double* __restrict__ s,
                                                  const double* __restrict__ rs, const int* __restrict__ status,
                                                  BOFluidView F, double* __restrict__ props) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const BOCell c = bo_cell(F, p[i], s[3 * (size_t)i], s[3 * (size_t)i + 2], rs[i], status[i]);
  store_cell(props + (size_t)i * kBOProps, c);
}

// Start-of-step accumulation [pv b_w s_w, pv b_o s_o, pv (b_g s_g + rs b_o s_o)]^0.
__global__ __launch_bounds__(256) void k_bo_acc0(int nc, const double* __restrict__ p0, const double* __restrict__ s0,
                                                 const double* __restrict__ rs0, const int* __restrict__ st0,
                                                 const double* __restrict__ pv, BOFluidView F,
                                                 double* __restrict__ acc0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const double p = p0[i], r = rs0[i];
  double d;
  const double pvm = pv[i] * quad(F.cr, p - F.pr_ref);
  const double bw = b_wat(F, p, d);
  const double bo = pvto(F, p, r, st0[i] == 3).b;
  const double bg = b_gas(F, p, d);
  const double sw = s0[3 * (size_t)i], so = s0[3 * (size_t)i + 1], sg = s0[3 * (size_t)i + 2];
  acc0[3 * (size_t)i] = pvm * bw * sw;
  acc0[3 * (size_t)i + 1] = pvm * bo * so;
  acc0[3 * (size_t)i + 2] = pvm * (bg * sg + r * bo * so);
}

// Face loop: LPR lanes per cell row walk the row's BCSR entries; each lane
// evaluates the three phase fluxes of one connection with derivatives in
// both cells (6 unknowns), writes the off-diagonal block and accumulates
// the diagonal; lane 0 adds the accumulation terms.
template <int LPR>
__global__ __launch_bounds__(256) void k_bo_cells(int nc, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, const int* __restrict__ diag,
                                                  const int* __restrict__ bconn, const double* __restrict__ bT,
                                                  const double* __restrict__ bG, const double* __restrict__ p,
                                                  const double* __restrict__ s, const int* __restrict__ status,
                                                  const double* __restrict__ props, const double* __restrict__ acc0,
                                                  const double* __restrict__ pv, BOFluidView F, double inv_dt,
                                                  double* __restrict__ r, double* __restrict__ J) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  const bool active = row < nc;
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = 0.0;
  if (active) {
    const double pr = p[row];
    const BOCell A = load_cell(props + (size_t)row * kBOProps);
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) {
      const int c = col[k];
      if (c == row) continue;
      const BOCell Bn = load_cell(props + (size_t)c * kBOProps);
      const double T = bT[k], G = bG[k];
      const bool isN0 = (bconn[k] & 1) == 0;
      const Dl<6> dP = dvar<6>(p[c], 3) - dvar<6>(pr, 0);
      Dl<6> Fq[3];
      Dl<6> Fo;
      bool up_o = true;
      // water, oil, gas: pot = dp - avg(rho) G, v = -T mob_up pot
#pragma unroll
      for (int ph = 0; ph < 3; ++ph) {
        const Dl<3>& br = ph == 0 ? A.bw : (ph == 1 ? A.bo : A.bg);
        const Dl<3>& bn = ph == 0 ? Bn.bw : (ph == 1 ? Bn.bo : Bn.bg);
        const Dl<3>& mr = ph == 0 ? A.mw : (ph == 1 ? A.mo : A.mg);
        const Dl<3>& mn = ph == 0 ? Bn.mw : (ph == 1 ? Bn.mo : Bn.mg);
        Dl<6> rho_r, rho_n;
        if (ph == 0) {
          rho_r = F.rhoW * lift<6, 0>(br);
          rho_n = F.rhoW * lift<6, 3>(bn);
        } else if (ph == 1) {
          rho_r = lift<6, 0>(br) * (F.rhoG * lift<6, 0>(A.rs) + dc<6>(F.rhoO));
          rho_n = lift<6, 3>(bn) * (F.rhoG * lift<6, 3>(Bn.rs) + dc<6>(F.rhoO));
        } else {
          rho_r = F.rhoG * lift<6, 0>(br);
          rho_n = F.rhoG * lift<6, 3>(bn);
        }
        const Dl<6> pot = dP - (0.5 * G) * (rho_r + rho_n);
        const bool up_r = isN0 ? (pot.v <= 0.0) : (pot.v < 0.0);
        const Dl<6> mob = up_r ? lift<6, 0>(mr) : lift<6, 3>(mn);
        const Dl<6> bu = up_r ? lift<6, 0>(br) : lift<6, 3>(bn);
        const Dl<6> v = (-T) * (mob * pot);
        Fq[ph] = bu * v;
        if (ph == 1) {
          Fo = Fq[1];
          up_o = up_r;
        }
      }
      const Dl<6> rsu = up_o ? lift<6, 0>(A.rs) : lift<6, 3>(Bn.rs);
      Fq[2] = Fq[2] + rsu * Fo;
      double* blk = J + (size_t)k * 9;
#pragma unroll
      for (int eq = 0; eq < 3; ++eq) {
        acc[eq] += Fq[eq].v;
#pragma unroll
        for (int v = 0; v < 3; ++v) {
          acc[3 + 3 * eq + v] += Fq[eq].d[v];
          blk[3 * eq + v] = Fq[eq].d[3 + v];
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 12; ++q) acc[q] = lane_sum(acc[q], LPR);
  if (active && lane == 0) {
    const BOCell A = load_cell(props + (size_t)row * kBOProps);
    const double pr = p[row];
    const int st = status[row];
    const Dl<3> P = dvar<3>(pr, 0), SW = dvar<3>(s[3 * (size_t)row], 1);
    const Dl<3> SG = st == 3 ? dvar<3>(s[3 * (size_t)row + 2], 2) : dc<3>(0.0);
    const Dl<3> SO = 1.0 - SW - SG;
    const double x = F.cr * (pr - F.pr_ref);
    const Dl<3> PV = chain(pv[row] * (1.0 + x + 0.5 * x * x), pv[row] * F.cr * (1.0 + x), P);
    Dl<3> R[3];
    R[0] = (PV * A.bw * SW - dc<3>(acc0[3 * (size_t)row])) * dc<3>(inv_dt);
    R[1] = (PV * A.bo * SO - dc<3>(acc0[3 * (size_t)row + 1])) * dc<3>(inv_dt);
    R[2] = (PV * (A.bg * SG + A.rs * A.bo * SO) - dc<3>(acc0[3 * (size_t)row + 2])) * dc<3>(inv_dt);
    double* blk = J + (size_t)diag[row] * 9;
#pragma unroll
    for (int eq = 0; eq < 3; ++eq) {
      r[3 * (size_t)row + eq] = R[eq].v + acc[eq];
#pragma unroll
      for (int v = 0; v < 3; ++v) blk[3 * eq + v] = R[eq].d[v] + acc[3 + 3 * eq + v];
    }
  }
}

// One 64-wide block per well (_well :383-436; with the facility's pieces,
// FacilityModel._single_well, facility_model.py:2434-2549): connection
// rates, well equations and the derivative pieces of the well model, with
// the wellbore mixture of the producing connections driving the injecting
// ones.  Duals carry (p, sW, x) of the connection cell and the well's bhp.
// cdp is the connection pressure drop below the bhp datum of every
// perforation (zero for the hand-rolled model).  The control row is
// ctrl = c_w qWs + c_o qOs + c_g qGs + c_b bhp - val (wctrl[4 w + .]); the
// three rate-closure rows are scaled by rsc[a] (the facility's rhoScale,
// one per phase; 1 for the hand-rolled model).  A shut well (wstat = 0)
// takes no flux: its rows are q = 0 and bhp = the mean of its perforations'
// pressures.  qvol receives every perforation's reservoir volume rates (a
// producing one's phase rates, an injecting one's total times compi).
__global__ __launch_bounds__(64) void k_bo_wells(
    int nw, const int* __restrict__ well_ptr, const int* __restrict__ perf_cell, const double* __restrict__ WI,
    const double* __restrict__ cdp, const double* __restrict__ wctrl, const int* __restrict__ wstat,
    const double* __restrict__ wval, const double* __restrict__ compi, const double* __restrict__ p,
    const double* __restrict__ props, const double* __restrict__ q, const double* __restrict__ bhp, double rsc0,
    double rsc1, double rsc2, size_t nc3, double* __restrict__ r, double* __restrict__ rq, double* __restrict__ dq,
    double* __restrict__ Bp, double* __restrict__ Cp, double* __restrict__ Um, double* __restrict__ V,
    double* __restrict__ Dp, int* __restrict__ mixflag, double* __restrict__ qvol) {
  const int w = blockIdx.x, lane = threadIdx.x;
  const int j0 = well_ptr[w], j1 = well_ptr[w + 1];
  const double bh = bhp[w];
  const double qs[3] = {q[3 * (size_t)w], q[3 * (size_t)w + 1], q[3 * (size_t)w + 2]};
  const double rsc[3] = {rsc0, rsc1, rsc2};
  if (wstat[w] == 0) {
    const double inv_n = j1 > j0 ? 1.0 / (double)(j1 - j0) : 0.0;
    double psum = 0.0;
    for (int j = j0 + lane; j < j1; j += 64) {
      psum += p[perf_cell[j]];
      for (int a = 0; a < 3; ++a) rq[3 * (size_t)j + a] = qvol[3 * (size_t)j + a] = 0.0;
      for (int k = 0; k < 9; ++k) dq[9 * (size_t)j + k] = Um[9 * (size_t)j + k] = V[9 * (size_t)j + k] = 0.0;
      for (int k = 0; k < 12; ++k) Bp[12 * (size_t)j + k] = Cp[12 * (size_t)j + k] = 0.0;
      Cp[12 * (size_t)j + 9] = -inv_n;      // d ctrl / d p of the connection cell
    }
    psum = lane_sum(psum, 64);
    if (lane == 0) {
      double* rw = r + nc3 + 4 * (size_t)w;
      for (int a = 0; a < 3; ++a) rw[a] = qs[a];
      rw[3] = bh - psum * inv_n;
      double* D = Dp + 16 * (size_t)w;
      for (int k = 0; k < 16; ++k) D[k] = (k % 5 == 0) ? 1.0 : 0.0;
      mixflag[w] = 0;
    }
    return;
  }
  // Phase A: surface rates of the producing connections and d/dbhp
  double S[3] = {0, 0, 0}, dSb[3] = {0, 0, 0};
  int anyprod = 0;
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const double dd = -(bh + cdp[j]) + p[c];
    if (dd < 0.0) continue;
    anyprod = 1;
    const BOCell C = load_cell(props + (size_t)c * kBOProps);
    const double t = WI[j];
    const double cw = -(t * C.mw.v) * dd, co = -(t * C.mo.v) * dd, cg = -(t * C.mg.v) * dd;
    S[0] += C.bw.v * cw;
    S[1] += C.bo.v * co;
    S[2] += C.bg.v * cg + C.rs.v * (C.bo.v * co);
    dSb[0] += C.bw.v * t * C.mw.v;
    dSb[1] += C.bo.v * t * C.mo.v;
    dSb[2] += C.bg.v * t * C.mg.v + C.rs.v * C.bo.v * t * C.mo.v;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    S[a] = lane_sum(S[a], 64);
    dSb[a] = lane_sum(dSb[a], 64);
  }
  anyprod = __any(anyprod) ? 1 : 0;
  // Phase B: wellbore mixture and its derivatives (redundant in every lane)
  double pos[3], wbq[3];
  for (int a = 0; a < 3; ++a) {
    pos[a] = qs[a] > 0.0 ? 1.0 : 0.0;
    wbq[a] = qs[a] * pos[a] - S[a];
  }
  const double wbqt = wbq[0] + wbq[1] + wbq[2];
  const bool dead = !anyprod || wbqt == 0.0;
  double mix[3], Am[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (dead) {
    for (int a = 0; a < 3; ++a) mix[a] = compi[3 * (size_t)w + a];
  } else {
    const double inv = 1.0 / wbqt;
    for (int a = 0; a < 3; ++a) {
      mix[a] = wbq[a] / wbqt;
      for (int b = 0; b < 3; ++b) Am[a][b] = (a == b ? inv : 0.0) - wbq[a] * inv * inv;
    }
  }
  // F = d mix / d (qWs, qOs, qGs, bhp)
  double Fm[3][4];
  for (int a = 0; a < 3; ++a) {
    double fb = 0.0;
    for (int b = 0; b < 3; ++b) {
      Fm[a][b] = Am[a][b] * pos[b];
      fb += Am[a][b] * dSb[b];
    }
    Fm[a][3] = -fb;
  }
  // Phase C: per-connection rates and derivative pieces
  double Q[3] = {0, 0, 0}, dQb[3] = {0, 0, 0};
  double E[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // sum_j Um_j
  int mix_used = 0;
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const BOCell C = load_cell(props + (size_t)c * kBOProps);
    const double t = WI[j];
    const Dl<4> dd = lift<4, 0>(dvar<3>(p[c], 0)) - dvar<4>(bh + cdp[j], 3);
    const Dl<4> bw = lift<4, 0>(C.bw), bo = lift<4, 0>(C.bo), bg = lift<4, 0>(C.bg), rs = lift<4, 0>(C.rs);
    Dl<4> cq[3];
    double U[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double Vj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (dd.v < 0.0) {
      const Dl<4> mt = lift<4, 0>(C.mw) + lift<4, 0>(C.mo) + lift<4, 0>(C.mg);
      const Dl<4> cqt = (-t) * (mt * dd);
      const Dl<4> vr = mix[0] / bw + mix[1] / bo + (dc<4>(mix[2]) - mix[1] * rs) / bg;
      const Dl<4> cis = cqt / vr;
      for (int a = 0; a < 3; ++a) cq[a] = mix[a] * cis;
      const double q2 = cqt.v / (vr.v * vr.v);
      const double dvr[3] = {1.0 / bw.v, 1.0 / bo.v - rs.v / bg.v, 1.0 / bg.v};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) U[a][b] = (a == b ? cis.v : 0.0) - mix[a] * q2 * dvr[b];
      mix_used = 1;
      for (int a = 0; a < 3; ++a) qvol[3 * (size_t)j + a] = cqt.v * compi[3 * (size_t)w + a];
    } else {
      const Dl<4> cw = (-t) * (lift<4, 0>(C.mw) * dd);
      const Dl<4> co = (-t) * (lift<4, 0>(C.mo) * dd);
      const Dl<4> cg = (-t) * (lift<4, 0>(C.mg) * dd);
      cq[0] = bw * cw;
      cq[1] = bo * co;
      cq[2] = bg * cg + rs * (bo * co);
      qvol[3 * (size_t)j] = cw.v;
      qvol[3 * (size_t)j + 1] = co.v;
      qvol[3 * (size_t)j + 2] = cg.v;
      // d mix / d (p, sW, x)_c through wbq = ... - sum cq_ps
      for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 3; ++k)
          Vj[a][k] = -(Am[a][0] * cq[0].d[k] + Am[a][1] * cq[1].d[k] + Am[a][2] * cq[2].d[k]);
    }
    double um[3][3];
    for (int a = 0; a < 3; ++a) {
      Q[a] += cq[a].v;
      dQb[a] += cq[a].d[3];
      rq[3 * (size_t)j + a] = cq[a].v;
      for (int k = 0; k < 3; ++k) {
        dq[9 * (size_t)j + 3 * a + k] = cq[a].d[k];
        um[a][k] = -U[a][k];
        E[a][k] += um[a][k];
        Um[9 * (size_t)j + 3 * a + k] = um[a][k];
        V[9 * (size_t)j + 3 * a + k] = Vj[a][k];
        Cp[12 * (size_t)j + 3 * a + k] = -cq[a].d[k];
      }
      for (int qq = 0; qq < 4; ++qq)
        Bp[12 * (size_t)j + 4 * a + qq] = (qq == 3 ? -cq[a].d[3] : 0.0) + um[a][0] * Fm[0][qq] +
                                          um[a][1] * Fm[1][qq] + um[a][2] * Fm[2][qq];
    }
    for (int k = 0; k < 3; ++k) Cp[12 * (size_t)j + 9 + k] = 0.0;
  }
  for (int a = 0; a < 3; ++a) {
    Q[a] = lane_sum(Q[a], 64);
    dQb[a] = lane_sum(dQb[a], 64);
    for (int b = 0; b < 3; ++b) E[a][b] = lane_sum(E[a][b], 64);
  }
  mix_used = __any(mix_used) && !dead ? 1 : 0;
  // Phase D: well rows (the mixture makes every connection's rate depend on
  // the producing cells: Cp += E V)
  for (int j = j0 + lane; j < j1; j += 64) {
    double* cp = Cp + 12 * (size_t)j;
    const double* vv = V + 9 * (size_t)j;
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 3; ++k)
        cp[3 * a + k] = rsc[a] * (cp[3 * a + k] + E[a][0] * vv[k] + E[a][1] * vv[3 + k] + E[a][2] * vv[6 + k]);
  }
  if (lane == 0) {
    const double* cc = wctrl + 4 * (size_t)w;
    double* rw = r + nc3 + 4 * (size_t)w;
    for (int a = 0; a < 3; ++a) rw[a] = rsc[a] * (qs[a] - Q[a]);
    rw[3] = cc[0] * qs[0] + cc[1] * qs[1] + cc[2] * qs[2] + cc[3] * bh - wval[w];
    double* D = Dp + 16 * (size_t)w;
    for (int a = 0; a < 3; ++a)
      for (int qq = 0; qq < 4; ++qq)
        D[4 * a + qq] = rsc[a] * ((a == qq ? 1.0 : 0.0) + (qq == 3 ? -dQb[a] : 0.0) + E[a][0] * Fm[0][qq] +
                                  E[a][1] * Fm[1][qq] + E[a][2] * Fm[2][qq]);
    for (int qq = 0; qq < 4; ++qq) D[12 + qq] = cc[qq];
    mixflag[w] = mix_used;
  }
}

// Subtract connection rates and their local derivatives from the cell rows.
__global__ void k_bo_perf_scatter(int npc, const int* __restrict__ pc_cell, const int* __restrict__ pc_ptr,
                                  const int* __restrict__ pc_perf, const int* __restrict__ diag,
                                  const double* __restrict__ rq, const double* __restrict__ dq,
                                  double* __restrict__ r, double* __restrict__ J) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npc) return;
  const int c = pc_cell[u];
  double* blk = J + (size_t)diag[c] * 9;
  for (int t = pc_ptr[u]; t < pc_ptr[u + 1]; ++t) {
    const int j = pc_perf[t];
    for (int a = 0; a < 3; ++a) r[3 * (size_t)c + a] -= rq[3 * (size_t)j + a];
    for (int q = 0; q < 9; ++q) blk[q] -= dq[9 * (size_t)j + q];
  }
}

// Convergence partials (_conv_scaling/_converged): max|r_a/pv(p)| (3),
// sum r_a (3), sum 1/b_a (3), sum p, sum rs, sum pv(p).
constexpr int kBOConvBlocks = 512;
constexpr int kBOConv = 12;
__global__ __launch_bounds__(256) void k_bo_conv(int nc, const double* __restrict__ r, const double* __restrict__ p,
                                                 const double* __restrict__ rs, const int* __restrict__ status,
                                                 const double* __restrict__ pv, BOFluidView F,
                                                 double* __restrict__ partial) {
  __shared__ double sh[kBOConv][4];
  const size_t chunk = ((size_t)nc + gridDim.x - 1) / gridDim.x;
  const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min((size_t)nc, i0 + chunk);
  double v[kBOConv];
  for (int q = 0; q < kBOConv; ++q) v[q] = 0.0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double pi = p[i];
    const double pvc = pv[i] * quad(F.cr, pi - F.pr_ref);
    double d;
    const double b[3] = {b_wat(F, pi, d), pvto(F, pi, rs[i], status[i] == 3).b, b_gas(F, pi, d)};
    for (int a = 0; a < 3; ++a) {
      const double ra = r[3 * i + a];
      v[a] = fmax(v[a], fabs(ra / pvc));
      v[3 + a] += ra;
      v[6 + a] += 1.0 / b[a];
    }
    v[9] += pi;
    v[10] += rs[i];
    v[11] += pvc;
  }
  for (int off = 32; off > 0; off >>= 1) {
    for (int q = 0; q < 3; ++q) v[q] = fmax(v[q], __shfl_xor(v[q], off, 64));
    for (int q = 3; q < kBOConv; ++q) v[q] += __shfl_xor(v[q], off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < kBOConv; ++q) sh[q][wid] = v[q];
  __syncthreads();
  if (threadIdx.x < kBOConv) {
    const int q = threadIdx.x;
    double a = sh[q][0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) a = q < 3 ? fmax(a, sh[q][k]) : a + sh[q][k];
    partial[(size_t)blockIdx.x * kBOConv + q] = a;
  }
}

// One wave per quantity: lanes stride over the block partials, then a
// shuffle tree (fixed order, deterministic); the first 3 quantities are maxima.
__global__ void k_bo_conv_final(int nblk, const double* __restrict__ partial, double* __restrict__ out) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (q >= kBOConv) return;
  const bool is_max = q < 3;
  double a = 0.0;   // the maxima are of absolute values
  for (int b = lane; b < nblk; b += 64) {
    const double v = partial[(size_t)b * kBOConv + q];
    a = is_max ? fmax(a, v) : a + v;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(a, off, 64);
    a = is_max ? fmax(a, o) : a + o;
  }
  if (lane == 0) out[q] = a;
}

// Newton update with the variable switch and the flash (update_state
// :439-470, _flash :472-509), one thread per cell; well unknowns after.
// rule 0 is the hand-rolled model's: each saturation change clipped to
// +-ds_max, pressure unlimited.  rule 1 is update_state_black_oil_generic's
// (update_state_black_oil_generic.py:30-80): the pressure change capped at
// dp_max_rel |p|, the rs change at drs_max, and the three saturation changes
// scaled together so that the largest is at most ds_max; the flash is the
// same (compute_flash_black_oil.py:22-140 without vaporised oil).
__global__ void k_bo_update(int nc, int nw, const double* __restrict__ dx, double ds_max, int rule,
                            double dp_max_rel, double drs_max, BOFluidView F, double* __restrict__ p,
                            double* __restrict__ s, double* __restrict__ rs, int* __restrict__ status,
                            double* __restrict__ q, double* __restrict__ bhp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) {
    const double etol = 1.4901161193847656e-08;   // sqrt(DBL_EPSILON)
    const bool st1 = status[i] == 1, st3 = status[i] == 3;
    const double p_prev = p[i], rs_prev = rs[i];
    const double sw_prev = s[3 * (size_t)i], so_prev = s[3 * (size_t)i + 1], sg_prev = s[3 * (size_t)i + 2];
    const double dp = dx[3 * (size_t)i], dsw = dx[3 * (size_t)i + 1], dxx = dx[3 * (size_t)i + 2];
    const double dsg = st3 ? dxx : 0.0;
    const double ds[3] = {dsw, -(dsg + dsw), dsg};
    double pn, r, sw, so, sg;
    if (rule == 0) {
      pn = p_prev + dp;
      r = fmax(rs_prev + (st1 ? dxx : 0.0), 0.0);
      sw = sw_prev + fmin(fmax(ds[0], -ds_max), ds_max);
      so = so_prev + fmin(fmax(ds[1], -ds_max), ds_max);
      sg = sg_prev + fmin(fmax(ds[2], -ds_max), ds_max);
    } else {
      const double cap = dp_max_rel * fabs(p_prev);
      pn = p_prev + fmin(fmax(dp, -cap), cap);
      r = fmax(rs_prev + fmin(fmax(st1 ? dxx : 0.0, -drs_max), drs_max), 0.0);
      const double largest = fmax(fmax(fabs(ds[0]), fabs(ds[1])), fabs(ds[2]));
      const double sc = largest > 0.0 ? fmin(ds_max / largest, 1.0) : 1.0;
      sw = sw_prev + ds[0] * sc;
      so = so_prev + ds[1] * sc;
      sg = sg_prev + ds[2] * sc;
    }
    // flash
    const bool wat_only = sw > 1.0 - etol;
    double d;
    const double rsat = rs_sat(F, pn, d);
    const double rsat0 = rs_sat(F, p_prev, d);
    bool gas = ((sg > 0.0) || (r == 0.0)) && !st1;
    gas = gas || wat_only;
    gas = gas || ((sg < 0.0) && (sg_prev > etol));
    const bool ix2 = (r > rsat * (1.0 + etol)) && st1 && (rs_prev > rsat0 * (1.0 - etol));
    if (ix2) sg = 0.0;
    gas = gas || ix2;
    if (sg < 0.0) {
      sw /= 1.0 - sg;
      so /= 1.0 - sg;
      sg = 0.0;
    }
    if (so < 0.0) {
      sw /= 1.0 - so;
      sg /= 1.0 - so;
      so = 0.0;
    }
    if (sw < 0.0) {
      so /= 1.0 - sw;
      sg /= 1.0 - sw;
      sw = 0.0;
    }
    r = gas ? rsat : fmin(rsat, r);
    r = fmax(r, 0.0);
    const double tot = sw + so + sg;
    p[i] = pn;
    rs[i] = r;
    status[i] = gas ? 3 : 1;
    s[3 * (size_t)i] = sw / tot;
    s[3 * (size_t)i + 1] = so / tot;
    s[3 * (size_t)i + 2] = sg / tot;
  } else if (i < nc + nw) {
    const int w = i - nc;
    const double* dw = dx + 3 * (size_t)nc + 4 * (size_t)w;
    for (int a = 0; a < 3; ++a) q[3 * (size_t)w + a] += dw[a];
    bhp[w] += dw[3];
  }
}

__global__ void k_bo_negate(size_t n, const double* __restrict__ x, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = -x[i];
}

// ---------------------------------------------------------------------------
// fluid tables (init_deck_blackoil_fluid, init_deck_adi_fluid.py:63-194)
// ---------------------------------------------------------------------------
void BOFluidTables::build(const double* sc, int n_pvdg, const double* pvdg, int n_key, const double* keys,
                          const int* branch_ptr, const double* branch, int n_swof, const double* swof, int n_sgof,
                          const double* sgof) {
  PMX_CHECK(n_pvdg >= 1 && n_key >= 2 && n_swof >= 2 && n_sgof >= 2, "black-oil tables are too short");
  tab.clear();
  bptr.clear();
  BOFluidView& L = layout;
  L = BOFluidView{};
  L.rhoW = sc[0];
  L.rhoO = sc[1];
  L.rhoG = sc[2];
  L.pw_ref = sc[3];
  L.bw_ref = sc[4];
  L.cw = sc[5];
  L.muw = sc[6];
  L.pr_ref = sc[7];
  L.cr = sc[8];
  auto put = [&](const std::vector<double>& v) {
    const int off = (int)tab.size();
    tab.insert(tab.end(), v.begin(), v.end());
    return off;
  };
  auto check_ascending = [](const std::vector<double>& x, const char* what) {
    for (size_t i = 1; i < x.size(); ++i)
      PMX_CHECK(x[i] > x[i - 1], std::string(what) + " must be strictly increasing");
  };
  // PVDG with a leading row at p = 0 carrying the first row's values
  {
    std::vector<double> pg(n_pvdg + 1), bg(n_pvdg + 1), mg(n_pvdg + 1);
    pg[0] = 0.0;
    bg[0] = 1.0 / pvdg[1];
    mg[0] = pvdg[2];
    for (int i = 0; i < n_pvdg; ++i) {
      pg[i + 1] = pvdg[3 * i];
      bg[i + 1] = 1.0 / pvdg[3 * i + 1];
      mg[i + 1] = pvdg[3 * i + 2];
    }
    check_ascending(pg, "PVDG pressures");
    L.n_g = n_pvdg + 1;
    L.g_p = put(pg);
    L.g_b = put(bg);
    L.g_mu = put(mg);
  }
  // PVTO: per rs key a branch of (p, Bo, mu_o) rows; a single-row branch
  // gets a second row at the next representable pressure
  {
    std::vector<double> key(keys, keys + n_key), satp(n_key), satb(n_key), satmu(n_key);
    check_ascending(key, "PVTO keys");
    std::vector<double> dp, bb, mu;
    bptr.push_back(0);
    for (int k = 0; k < n_key; ++k) {
      const int a = branch_ptr[k], b = branch_ptr[k + 1];
      PMX_CHECK(b > a, "empty PVTO branch");
      std::vector<double> px, bx, mx;
      for (int i = a; i < b; ++i) {
        px.push_back(branch[3 * i]);
        bx.push_back(1.0 / branch[3 * i + 1]);
        mx.push_back(branch[3 * i + 2]);
      }
      if (px.size() == 1) {
        px.push_back(std::nextafter(px[0], INFINITY));
        bx.push_back(bx[0]);
        mx.push_back(mx[0]);
      }
      check_ascending(px, "PVTO branch pressures");
      satp[k] = px[0];
      satb[k] = bx[0];
      satmu[k] = mx[0];
      for (size_t i = 0; i < px.size(); ++i) {
        dp.push_back(px[i] - px[0]);
        bb.push_back(bx[i]);
        mu.push_back(mx[i]);
      }
      bptr.push_back((int)dp.size());
    }
    check_ascending(satp, "PVTO saturation pressures");
    L.nkey = n_key;
    L.key = put(key);
    L.sat_p = put(satp);
    L.sat_b = put(satb);
    L.sat_mu = put(satmu);
    L.br_dp = put(dp);
    L.br_b = put(bb);
    L.br_mu = put(mu);
    std::vector<double> rp(n_key + 1), rk(n_key + 1);
    rp[0] = 0.0;
    rk[0] = 0.0;
    for (int k = 0; k < n_key; ++k) {
      rp[k + 1] = satp[k];
      rk[k + 1] = key[k];
    }
    check_ascending(rp, "saturation pressures");
    L.n_rs = n_key + 1;
    L.rs_p = put(rp);
    L.rs_k = put(rk);
  }
  // SWOF / SGOF extended with flat rows one unit beyond each end
  auto extend = [](const double* t4, int n, int col, bool x) {
    std::vector<double> v(n + 2);
    for (int i = 0; i < n; ++i) v[i + 1] = t4[4 * i + col];
    v[0] = t4[col] - (x ? 1.0 : 0.0);
    v[n + 1] = t4[4 * (n - 1) + col] + (x ? 1.0 : 0.0);
    return v;
  };
  L.swcon = swof[0];
  {
    const std::vector<double> s = extend(swof, n_swof, 0, true), krw = extend(swof, n_swof, 1, false),
                              krow = extend(swof, n_swof, 2, false);
    check_ascending(s, "SWOF saturations");
    const int n = n_swof + 2;
    std::vector<double> so(n), ko(n);
    for (int i = 0; i < n; ++i) {
      so[i] = 1.0 - s[n - 1 - i];
      ko[i] = krow[n - 1 - i];
    }
    L.n_w = n;
    L.w_s = put(s);
    L.w_kr = put(krw);
    L.n_ow = n;
    L.ow_s = put(so);
    L.ow_kr = put(ko);
  }
  {
    const std::vector<double> s = extend(sgof, n_sgof, 0, true), krg = extend(sgof, n_sgof, 1, false),
                              krog = extend(sgof, n_sgof, 2, false);
    check_ascending(s, "SGOF saturations");
    const int n = n_sgof + 2;
    std::vector<double> so(n), ko(n);
    for (int i = 0; i < n; ++i) {
      so[i] = 1.0 - (s[n - 1 - i] + L.swcon);
      ko[i] = krog[n - 1 - i];
    }
    L.n_gg = n;
    L.gg_s = put(s);
    L.gg_kr = put(krg);
    L.n_og = n;
    L.og_s = put(so);
    L.og_kr = put(ko);
  }
}

void BOFluidTables::upload(hipStream_t s) {
  dtab.upload(tab, s);
  dbptr.upload(bptr, s);
}

// ---------------------------------------------------------------------------
BOModelDev::BOModelDev(hipStream_t s) : stream(s) {
  S.B = 3;
  S.M = 4;
  S.X = 3;
  cpr = cpr_create();
}
BOModelDev::~BOModelDev() { cpr_destroy(cpr); }

void BOModelDev::set_topology(int nc_, int nconn_, const int* N, const double* T, const double* gdz,
                              const double* pv_host) {
  nc = nc_;
  nconn = nconn_;
  HostPattern pat;
  std::vector<int> econn;
  tpfa_pattern(nc, nconn, N, pat, econn);
  S.set_pattern(pat, stream);
  const int nnz = pat.rowptr[nc];
  std::vector<double> hT(nnz, 0.0), hG(nnz, 0.0);
  for (int k = 0; k < nnz; ++k) {
    const int code = econn[k];
    if (code < 0) continue;
    const int f = code >> 1;
    hT[k] = T[f];
    hG[k] = (code & 1) ? -gdz[f] : gdz[f];
  }
  bconn.upload(econn, stream);
  bT.upload(hT, stream);
  bG.upload(hG, stream);
  pv.upload(pv_host, nc, stream);
  props.alloc((size_t)nc * kBOProps);
  acc0.alloc((size_t)nc * 3);
  w_cpr.alloc(3);
  conv_partial.alloc((size_t)kBOConvBlocks * kBOConv + 16);
  hier_weights = hT;
  PMX_HIP(hipStreamSynchronize(stream));
}

void BOModelDev::set_wells(int nw_, const int* perf_ptr, const int* perf_cells, const double* WI_h,
                           const int* types, const double* vals, const double* compi_h) {
  nw = nw_;
  std::vector<int> ptr(perf_ptr, perf_ptr + nw + 1);
  const int np = nw ? ptr[nw] : 0;
  std::vector<int> cells(perf_cells, perf_cells + np);
  for (int c : cells) PMX_CHECK(c >= 0 && c < nc, "well connection cell out of range");
  S.set_wells(nw, ptr, cells, stream);
  WI.upload(WI_h, np, stream);
  wtype.upload(types, nw, stream);
  wval.upload(vals, nw, stream);
  compi.upload(compi_h, 3 * (size_t)nw, stream);
  rq.alloc(3 * (size_t)std::max(np, 1));
  dq.alloc(9 * (size_t)std::max(np, 1));
  qvol.alloc(3 * (size_t)std::max(np, 1));
  mixflag.alloc(std::max(nw, 1));
  // the hand-rolled controls: bhp - val, or qWs + qOs + qGs - val
  std::vector<double> coef(4 * (size_t)std::max(nw, 1), 0.0);
  for (int k = 0; k < nw; ++k) {
    if (types[k] == 0) coef[4 * k + 3] = 1.0;
    else coef[4 * k] = coef[4 * k + 1] = coef[4 * k + 2] = 1.0;
  }
  wctrl.upload(coef.data(), coef.size(), stream);
  wstat.upload(std::vector<int>(std::max(nw, 1), 1).data(), std::max(nw, 1), stream);
  cdp.upload(std::vector<double>(std::max(np, 1), 0.0).data(), std::max(np, 1), stream);
  PMX_HIP(hipStreamSynchronize(stream));
}

void BOModelDev::set_controls(const double* coef, const double* val, const int* status, const double* rscale) {
  if (nw > 0) {
    wctrl.upload(coef, 4 * (size_t)nw, stream);
    wval.upload(val, nw, stream);
    wstat.upload(status, nw, stream);
  }
  for (int a = 0; a < 3; ++a) rsc[a] = rscale ? rscale[a] : 1.0;
  PMX_HIP(hipStreamSynchronize(stream));
}

void BOModelDev::set_connection_drops(const double* cdp_host) {
  const int np = nw ? S.nperf : 0;
  if (np > 0) cdp.upload(cdp_host, np, stream);
  PMX_HIP(hipStreamSynchronize(stream));
}

void BOModelDev::perf_volume_rates(double* out_host) {
  const int np = nw ? S.nperf : 0;
  if (np > 0) PMX_HIP(hipMemcpyAsync(out_host, qvol.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost, stream));
  PMX_HIP(hipStreamSynchronize(stream));
}

void BOModelDev::set_update_rule(int rule, double dpmr, double drsm) {
  PMX_CHECK(rule == 0 || rule == 1, "update rule is 0 (hand-rolled) or 1 (generic)");
  upd_rule = rule;
  dp_max_rel = dpmr;
  drs_max = drsm;
}

// Phase densities (water b_w rhoS_w, oil b_o (rhoS_o + rs rhoS_g), gas
// b_g rhoS_g) and mobilities of every cell at a state.
__global__ __launch_bounds__(256) void k_bo_densities(int nc, const double* __restrict__ p,
                                                      const double* __restrict__ s, const double* __restrict__ rs,
                                                      const int* __restrict__ status, BOFluidView F,
                                                      double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const BOCell c = bo_cell(F, p[i], s[3 * (size_t)i], s[3 * (size_t)i + 2], rs[i], status[i]);
  double* o = out + 6 * (size_t)i;
  o[0] = c.bw.v * F.rhoW;
  o[1] = c.bo.v * (F.rhoO + c.rs.v * F.rhoG);
  o[2] = c.bg.v * F.rhoG;
  o[3] = c.mw.v;
  o[4] = c.mo.v;
  o[5] = c.mg.v;
}

void BOModelDev::cell_densities(const double* p, const double* s, const double* rs, const int* status,
                                double* out) {
  PMX_CHECK(nc > 0, "topology not set");
  hipLaunchKernelGGL(k_bo_densities, dim3(grid_for(nc)), dim3(256), 0, stream, nc, p, s, rs, status, fluid.view(),
                     out);
  PMX_HIP(hipGetLastError());
}

void BOModelDev::begin_step(const double* p0, const double* s0, const double* rs0, const int* status0) {
  PMX_CHECK(nc > 0, "topology not set");
  hipLaunchKernelGGL(k_bo_acc0, dim3(grid_for(nc)), dim3(256), 0, stream, nc, p0, s0, rs0, status0, pv.p,
                     fluid.view(), acc0.p);
  PMX_HIP(hipGetLastError());
}

void BOModelDev::assemble(const double* p, const double* s, const double* rs, const int* status, const double* q,
                          const double* bhp, double dt) {
  PMX_CHECK(nc > 0, "topology not set");
  r.alloc(S.n());
  const BOFluidView F = fluid.view();
  hipLaunchKernelGGL(k_bo_props, dim3(grid_for(nc)), dim3(256), 0, stream, nc, p, s, rs, status, F, props.p);
  PMX_HIP(hipGetLastError());
  constexpr int LPR = 8;
  hipLaunchKernelGGL((k_bo_cells<LPR>), dim3(grid_for((size_t)nc * LPR)), dim3(256), 0, stream, nc, S.rowptr.p,
                     S.col.p, S.diag.p, bconn.p, bT.p, bG.p, p, s, status, props.p, acc0.p, pv.p, F, 1.0 / dt, r.p,
                     S.val.p);
  PMX_HIP(hipGetLastError());
  if (nw > 0) {
    hipLaunchKernelGGL(k_bo_wells, dim3(nw), dim3(64), 0, stream, nw, S.well_ptr.p, S.perf_cell.p, WI.p, cdp.p,
                       wctrl.p, wstat.p, wval.p, compi.p, p, props.p, q, bhp, rsc[0], rsc[1], rsc[2], (size_t)nc * 3,
                       r.p, rq.p, dq.p, S.Bp.p, S.Cp.p, S.Um.p, S.V.p, S.Dp.p, mixflag.p, qvol.p);
    PMX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bo_perf_scatter, dim3(grid_for(S.npc)), dim3(256), 0, stream, S.npc, S.pc_cell.p,
                       S.pc_ptr.p, S.pc_perf.p, S.diag.p, rq.p, dq.p, r.p, S.val.p);
    PMX_HIP(hipGetLastError());
    std::vector<int> mf(nw);
    PMX_HIP(hipMemcpyAsync(mf.data(), mixflag.p, sizeof(int) * nw, hipMemcpyDeviceToHost, stream));
    PMX_HIP(hipStreamSynchronize(stream));
    S.has_mix = false;
    for (int v : mf) S.has_mix |= (v != 0);
  }
}

void BOModelDev::convergence(const double* p, const double* rs, const int* status, double* out12,
                             double* wells_out) {
  hipLaunchKernelGGL(k_bo_conv, dim3(kBOConvBlocks), dim3(256), 0, stream, nc, r.p, p, rs, status, pv.p,
                     fluid.view(), conv_partial.p);
  PMX_HIP(hipGetLastError());
  double* fin = conv_partial.p + (size_t)kBOConvBlocks * kBOConv;
  hipLaunchKernelGGL(k_bo_conv_final, dim3(1), dim3(64 * kBOConv), 0, stream, kBOConvBlocks, conv_partial.p, fin);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipMemcpyAsync(out12, fin, sizeof(double) * kBOConv, hipMemcpyDeviceToHost, stream));
  if (nw > 0 && wells_out)
    PMX_HIP(hipMemcpyAsync(wells_out, r.p + 3 * (size_t)nc, sizeof(double) * 4 * nw, hipMemcpyDeviceToHost,
                           stream));
  PMX_HIP(hipStreamSynchronize(stream));
}

int BOModelDev::solve(const LinearSolverOptions& o, double* dx, SolveInfo& info, double pmean, double rsmean) {
  auto t0 = std::chrono::steady_clock::now();
  if (!cpr_has_hierarchy(cpr)) cpr_build_hierarchy(cpr, S, hier_weights, o, stream);
  // Pressure weights at the mean state: w = (1/b_w, 1/b_o - rs/b_g, 1/b_g)
  // makes the weighted accumulation pv (s_w + s_o + s_g) = pv independent of
  // the saturations and of the dissolved gas (the reference's 1/b_a weights,
  // linear_solvers.py:155-190, with the dissolved-gas term of the oil
  // column); b_o on the saturated curve.
  const BOFluidView H = fluid.host_view();
  double d;
  const double bw = b_wat(H, pmean, d);
  const double bo = pvto(H, pmean, rsmean, true).b;
  const double bg = b_gas(H, pmean, d);
  double wo = 1.0 / bo - rsmean / bg;
  if (!(wo > 0.1 / bo)) wo = 1.0 / bo;
  const double w[3] = {1.0 / bw, wo, 1.0 / bg};
  PMX_HIP(hipMemcpyAsync(w_cpr.p, w, sizeof(w), hipMemcpyHostToDevice, stream));
  cpr_setup(cpr, S, w_cpr.p, o, stream);
  PMX_HIP(hipStreamSynchronize(stream));
  const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rhs.alloc(S.n());
  hipLaunchKernelGGL(k_bo_negate, dim3(grid_for(S.n())), dim3(256), 0, stream, S.n(), r.p, rhs.p);
  PMX_HIP(hipGetLastError());
  const int st = cpr_solve(cpr, S, rhs.p, dx, o, info, stream);
  info.setup_ms = setup_ms;
  return st;
}

void BOModelDev::update(const double* dx, double ds_max, double* p, double* s, double* rs, int* status, double* q,
                        double* bhp) {
  hipLaunchKernelGGL(k_bo_update, dim3(grid_for((size_t)nc + nw)), dim3(256), 0, stream, nc, nw, dx, ds_max,
                     upd_rule, dp_max_rel, drs_max, fluid.view(), p, s, rs, status, q, bhp);
  PMX_HIP(hipGetLastError());
}

}  // namespace pmx
