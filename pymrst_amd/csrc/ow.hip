// Fully-implicit oil-water residual and Jacobian on the device.
//
// Equations (pymrst/modules/ad_blackoil/models/two_phase_oil_water_model.py
// assemble :493-546, with the fluid of init_simple_adi_fluid.py:220-239):
//   R_a(c) = (pv pvm(p) b_a(p) s_a - pv pvm(p0) b_a(p0) s_a0)/dt
//            + sum_faces div( upw(b_a) v_a ) - sum_perfs cq_a
//   v_a    = -T upw(mob_a) (dp - avg(rho_a) gdz),  upwind flag: pot <= 0 -> N0
// Well model: _well_contributions_batched (:355-491).  Derivatives are
// written out by hand; the forward-mode AD of the reference defines them.
//
// Layout: cell-major interleaved unknowns (p_c, sW_c); well-major well
// unknowns (qWs_w, qOs_w, bhp_w); residual rows (water_c, oil_c) and
// (eq_w, eq_o, ctrl) per well.  2x2 blocks are row-major:
//   [d water/dp, d water/dsW; d oil/dp, d oil/dsW].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <mutex>

#include "ow.h"
#include "ow_props.h"

namespace pmx {

template <int LPR>
__global__ __launch_bounds__(256) void k_ow_cells(int nc, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col, const int* __restrict__ diag,
                                                  const int* __restrict__ bconn, const double* __restrict__ bT,
                                                  const double* __restrict__ bG, const double* __restrict__ p,
                                                  const double* __restrict__ s, const double* __restrict__ p0,
                                                  const double* __restrict__ s0, const double* __restrict__ pv,
                                                  OWFluid f, double inv_dt, const double* __restrict__ dt_cell,
                                                  double* __restrict__ r, double* __restrict__ J) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  const bool active = row < nc;
  double fw = 0, fo = 0, d00 = 0, d01 = 0, d10 = 0, d11 = 0;
  double pr = 0, sr = 0;
  CellProps P{};
  if (active) {
    pr = p[row];
    sr = s[row];
    P = cell_props(f, pr, sr);
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) {
      const int c = col[k];
      if (c == row) continue;
      const double pn = p[c], sn = s[c];
      const CellProps Q = cell_props(f, pn, sn);
      const double T = bT[k], G = bG[k];
      const bool isN0 = (bconn[k] & 1) == 0;
      double Cw, wpr, wsr, wpn, wsn, Co, opr, osr, opn, osn;
      phase_flux(T, G, isN0, pr, pn, f.rho_w, P.bw, P.dbw, P.mw, P.dmw, Q.bw, Q.dbw, Q.mw, Q.dmw, Cw, wpr, wsr,
                 wpn, wsn);
      phase_flux(T, G, isN0, pr, pn, f.rho_o, P.bo, P.dbo, P.mo, P.dmo, Q.bo, Q.dbo, Q.mo, Q.dmo, Co, opr, osr,
                 opn, osn);
      fw += Cw;
      fo += Co;
      d00 += wpr;
      d01 += wsr;
      d10 += opr;
      d11 += osr;
      double* blk = J + (size_t)k * 4;
      blk[0] = wpn;
      blk[1] = wsn;
      blk[2] = opn;
      blk[3] = osn;
    }
  }
  fw = lane_sum(fw, LPR);
  fo = lane_sum(fo, LPR);
  d00 = lane_sum(d00, LPR);
  d01 = lane_sum(d01, LPR);
  d10 = lane_sum(d10, LPR);
  d11 = lane_sum(d11, LPR);
  if (active && lane == 0) {
    const double v = pv[row];
    const double pp0 = p0[row], ss0 = s0[row];
    double bw0, dbw0, bo0, dbo0;
    shrink(f.c_w, pp0, f.p_ref, bw0, dbw0);
    shrink(f.c_o, pp0, f.p_ref, bo0, dbo0);
    double m = 1.0, dm = 0.0, m0 = 1.0;
    if (f.has_cr) {
      m = 1.0 + f.c_r * (pr - f.p_ref);
      dm = f.c_r;
      m0 = 1.0 + f.c_r * (pp0 - f.p_ref);
    }
    const double pvm = v * m, pvm0 = v * m0;
    const double so = 1.0 - sr;
    if (dt_cell) inv_dt = 1.0 / dt_cell[row];   // per-cell step (ensemble members)
    const double aw = (pvm * P.bw * sr - pvm0 * bw0 * ss0) * inv_dt;
    const double ao = (pvm * P.bo * so - pvm0 * bo0 * (1.0 - ss0)) * inv_dt;
    r[2 * (size_t)row] = aw + fw;
    r[2 * (size_t)row + 1] = ao + fo;
    double* blk = J + (size_t)diag[row] * 4;
    blk[0] = v * (dm * P.bw + m * P.dbw) * sr * inv_dt + d00;
    blk[1] = pvm * P.bw * inv_dt + d01;
    blk[2] = v * (dm * P.bo + m * P.dbo) * so * inv_dt + d10;
    blk[3] = -pvm * P.bo * inv_dt + d11;
  }
}

// One 64-wide block per well: connection rates, well equations and all
// derivative pieces of the well model (mixture terms included).  The control
// row is ctrl = c_w qWs + c_o qOs + c_b bhp - val (wctrl[3 w + .] = c_w, c_o,
// c_b: a bhp target, a total or single-phase rate, or any linear form of the
// well's own unknowns); the two rate-closure rows are scaled by rs_w and
// rs_o.  A shut well (wstat = 0) takes no flux: its rows are q = 0 and bhp =
// the mean of its perforations' pressures.  qvol receives each perforation's
// reservoir volume rates (producing: the phase rates; injecting: the total
// times compi).
__global__ __launch_bounds__(64) void k_ow_wells(
    int nw, const int* __restrict__ well_ptr, const int* __restrict__ perf_cell, const double* __restrict__ WI,
    const double* __restrict__ cdp, const double* __restrict__ wctrl, const int* __restrict__ wstat,
    const double* __restrict__ wval, const double* __restrict__ compi, const double* __restrict__ p,
    const double* __restrict__ s, const double* __restrict__ qws, const double* __restrict__ qos,
    const double* __restrict__ bhp, OWFluid f, double rs_w, double rs_o, size_t nc2, double* __restrict__ r,
    double* __restrict__ rq, double* __restrict__ dq, double* __restrict__ Bp, double* __restrict__ Cp,
    double* __restrict__ Um, double* __restrict__ V, double* __restrict__ Dp, int* __restrict__ mixflag,
    double* __restrict__ qvol) {
  const int w = blockIdx.x, lane = threadIdx.x;
  const int j0 = well_ptr[w], j1 = well_ptr[w + 1];
  const double bh = bhp[w], qw = qws[w], qo = qos[w];
  if (wstat[w] == 0) {
    const double inv_n = j1 > j0 ? 1.0 / (double)(j1 - j0) : 0.0;
    double psum = 0.0;
    for (int j = j0 + lane; j < j1; j += 64) {
      psum += p[perf_cell[j]];
      rq[2 * (size_t)j] = rq[2 * (size_t)j + 1] = 0.0;
      qvol[2 * (size_t)j] = qvol[2 * (size_t)j + 1] = 0.0;
      for (int q = 0; q < 4; ++q) dq[4 * (size_t)j + q] = Um[4 * (size_t)j + q] = V[4 * (size_t)j + q] = 0.0;
      for (int q = 0; q < 6; ++q) Bp[6 * (size_t)j + q] = Cp[6 * (size_t)j + q] = 0.0;
      Cp[6 * (size_t)j + 4] = -inv_n;
    }
    psum = lane_sum(psum, 64);
    if (lane == 0) {
      r[nc2 + 3 * (size_t)w + 0] = qw;
      r[nc2 + 3 * (size_t)w + 1] = qo;
      r[nc2 + 3 * (size_t)w + 2] = bh - psum * inv_n;
      double* D = Dp + 9 * (size_t)w;
      for (int q = 0; q < 9; ++q) D[q] = (q % 4 == 0) ? 1.0 : 0.0;
      mixflag[w] = 0;
    }
    return;
  }
  // Phase A: producing connections
  double Sw = 0, So = 0, dSw_b = 0, dSo_b = 0;
  int anyprod = 0;
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const double pc = p[c];
    const double dd = -(bh + cdp[j]) + pc;
    if (dd < 0.0) continue;
    anyprod = 1;
    const CellProps P = cell_props(f, pc, s[c]);
    const double t = WI[j];
    const double cpw = -t * P.mw * dd, cpo = -t * P.mo * dd;
    Sw += P.bw * cpw;
    So += P.bo * cpo;
    dSw_b += P.bw * t * P.mw;
    dSo_b += P.bo * t * P.mo;
  }
  Sw = lane_sum(Sw, 64);
  So = lane_sum(So, 64);
  dSw_b = lane_sum(dSw_b, 64);
  dSo_b = lane_sum(dSo_b, 64);
  anyprod = __any(anyprod) ? 1 : 0;
  // Phase B: wellbore mixture (redundant in every lane)
  const double pw = qw > 0.0 ? 1.0 : 0.0, po = qo > 0.0 ? 1.0 : 0.0;
  const double wbq_w = qw * pw - Sw, wbq_o = qo * po - So;
  const double wbqt = wbq_w + wbq_o;
  const bool dead = (wbqt == 0.0) || !anyprod;
  double mixw, mixo, aww = 0, awo = 0, aow = 0, aoo = 0;
  if (dead) {
    mixw = compi[2 * w];
    mixo = compi[2 * w + 1];
  } else {
    const double inv = 1.0 / wbqt;
    mixw = wbq_w * inv;
    mixo = wbq_o * inv;
    aww = inv - wbq_w * inv * inv;
    awo = -wbq_w * inv * inv;
    aow = -wbq_o * inv * inv;
    aoo = inv - wbq_o * inv * inv;
  }
  // d mix / d (qw, qo, bhp)
  const double F[2][3] = {{aww * pw, awo * po, -(aww * dSw_b + awo * dSo_b)},
                          {aow * pw, aoo * po, -(aow * dSw_b + aoo * dSo_b)}};
  // Phase C: per-connection rates and derivative pieces
  double Qw = 0, Qo = 0, dQw_b = 0, dQo_b = 0;
  double E[2][2] = {{0, 0}, {0, 0}};   // sum_j Um_j  (d eq / d mix)
  int mix_used = 0;
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const double pc = p[c];
    const double dd = -(bh + cdp[j]) + pc;
    const CellProps P = cell_props(f, pc, s[c]);
    const double t = WI[j];
    double cw, co, lw[3], lo[3];   // rates and local d/d(p, s, bhp)
    double U[2][2] = {{0, 0}, {0, 0}};
    double Vj[2][2] = {{0, 0}, {0, 0}};
    if (dd < 0.0) {
      const double mt = P.mw + P.mo, dmt = P.dmw + P.dmo;
      const double cqt = -t * (mt * dd);
      const double g_p = -t * mt, g_s = -t * dmt * dd, g_b = t * mt;
      const double vr = mixw / P.bw + mixo / P.bo;
      const double dvr_p = -mixw * P.dbw / (P.bw * P.bw) - mixo * P.dbo / (P.bo * P.bo);
      const double cis = cqt / vr;
      const double q2 = cqt / (vr * vr);
      const double dcis_p = g_p / vr - q2 * dvr_p, dcis_s = g_s / vr, dcis_b = g_b / vr;
      cw = mixw * cis;
      co = mixo * cis;
      lw[0] = mixw * dcis_p; lw[1] = mixw * dcis_s; lw[2] = mixw * dcis_b;
      lo[0] = mixo * dcis_p; lo[1] = mixo * dcis_s; lo[2] = mixo * dcis_b;
      U[0][0] = cis - mixw * q2 / P.bw;
      U[0][1] = -mixw * q2 / P.bo;
      U[1][0] = -mixo * q2 / P.bw;
      U[1][1] = cis - mixo * q2 / P.bo;
      mix_used = 1;
    } else {
      const double cpw = -t * P.mw * dd, cpo = -t * P.mo * dd;
      cw = P.bw * cpw;
      co = P.bo * cpo;
      lw[0] = P.bw * (-t * P.mw) + P.dbw * cpw;
      lw[1] = P.bw * (-t * P.dmw * dd);
      lw[2] = P.bw * t * P.mw;
      lo[0] = P.bo * (-t * P.mo) + P.dbo * cpo;
      lo[1] = P.bo * (-t * P.dmo * dd);
      lo[2] = P.bo * t * P.mo;
      // d mix / d (p_c, s_c) through wbq = ... - sum cq_p
      Vj[0][0] = -(aww * lw[0] + awo * lo[0]);
      Vj[0][1] = -(aww * lw[1] + awo * lo[1]);
      Vj[1][0] = -(aow * lw[0] + aoo * lo[0]);
      Vj[1][1] = -(aow * lw[1] + aoo * lo[1]);
    }
    Qw += cw;
    Qo += co;
    dQw_b += lw[2];
    dQo_b += lo[2];
    rq[2 * (size_t)j] = cw;
    rq[2 * (size_t)j + 1] = co;
    if (dd < 0.0) {
      const double cqt = -t * ((P.mw + P.mo) * dd);
      qvol[2 * (size_t)j] = cqt * compi[2 * w];
      qvol[2 * (size_t)j + 1] = cqt * compi[2 * w + 1];
    } else {
      qvol[2 * (size_t)j] = -t * P.mw * dd;
      qvol[2 * (size_t)j + 1] = -t * P.mo * dd;
    }
    dq[4 * (size_t)j + 0] = lw[0];
    dq[4 * (size_t)j + 1] = lw[1];
    dq[4 * (size_t)j + 2] = lo[0];
    dq[4 * (size_t)j + 3] = lo[1];
    // Um = -U ; Bp = -(local d/d wellvars) + Um F
    double* um = Um + 4 * (size_t)j;
    um[0] = -U[0][0]; um[1] = -U[0][1]; um[2] = -U[1][0]; um[3] = -U[1][1];
    E[0][0] += um[0]; E[0][1] += um[1]; E[1][0] += um[2]; E[1][1] += um[3];
    double* bp = Bp + 6 * (size_t)j;
    for (int q = 0; q < 3; ++q) {
      bp[q] = (q == 2 ? -lw[2] : 0.0) + um[0] * F[0][q] + um[1] * F[1][q];
      bp[3 + q] = (q == 2 ? -lo[2] : 0.0) + um[2] * F[0][q] + um[3] * F[1][q];
    }
    double* vv = V + 4 * (size_t)j;
    vv[0] = Vj[0][0]; vv[1] = Vj[0][1]; vv[2] = Vj[1][0]; vv[3] = Vj[1][1];
    // local well-row derivatives (completed with E V below)
    double* cp = Cp + 6 * (size_t)j;
    cp[0] = -lw[0]; cp[1] = -lw[1];
    cp[2] = -lo[0]; cp[3] = -lo[1];
    cp[4] = 0.0;    cp[5] = 0.0;
  }
  Qw = lane_sum(Qw, 64);
  Qo = lane_sum(Qo, 64);
  dQw_b = lane_sum(dQw_b, 64);
  dQo_b = lane_sum(dQo_b, 64);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) E[a][b] = lane_sum(E[a][b], 64);
  mix_used = __any(mix_used) && !dead ? 1 : 0;
  // Phase D: well rows (the rate rows scaled)
  for (int j = j0 + lane; j < j1; j += 64) {
    double* cp = Cp + 6 * (size_t)j;
    const double* vv = V + 4 * (size_t)j;
    cp[0] = rs_w * (cp[0] + E[0][0] * vv[0] + E[0][1] * vv[2]);
    cp[1] = rs_w * (cp[1] + E[0][0] * vv[1] + E[0][1] * vv[3]);
    cp[2] = rs_o * (cp[2] + E[1][0] * vv[0] + E[1][1] * vv[2]);
    cp[3] = rs_o * (cp[3] + E[1][0] * vv[1] + E[1][1] * vv[3]);
  }
  if (lane == 0) {
    const double cw_ = wctrl[3 * w], co_ = wctrl[3 * w + 1], cb_ = wctrl[3 * w + 2];
    r[nc2 + 3 * (size_t)w + 0] = rs_w * (qw - Qw);
    r[nc2 + 3 * (size_t)w + 1] = rs_o * (qo - Qo);
    r[nc2 + 3 * (size_t)w + 2] = cw_ * qw + co_ * qo + cb_ * bh - wval[w];
    double* D = Dp + 9 * (size_t)w;
    const double loc[3][3] = {{1.0, 0.0, -dQw_b}, {0.0, 1.0, -dQo_b}, {cw_, co_, cb_}};
    const double rs[2] = {rs_w, rs_o};
    for (int a = 0; a < 3; ++a)
      for (int q = 0; q < 3; ++q)
        D[a * 3 + q] = a < 2 ? rs[a] * (loc[a][q] + E[a][0] * F[0][q] + E[a][1] * F[1][q]) : loc[a][q];
    mixflag[w] = mix_used;
  }
}

// Subtract connection rates and their local derivatives from the cell rows
// (unique perforated cells; connections in input order).
__global__ void k_ow_perf_scatter(int npc, const int* __restrict__ pc_cell, const int* __restrict__ pc_ptr,
                                  const int* __restrict__ pc_perf, const int* __restrict__ diag,
                                  const double* __restrict__ rq, const double* __restrict__ dq,
                                  double* __restrict__ r, double* __restrict__ J) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npc) return;
  const int c = pc_cell[u];
  double* blk = J + (size_t)diag[c] * 4;
  for (int t = pc_ptr[u]; t < pc_ptr[u + 1]; ++t) {
    const int j = pc_perf[t];
    r[2 * (size_t)c] -= rq[2 * (size_t)j];
    r[2 * (size_t)c + 1] -= rq[2 * (size_t)j + 1];
    for (int q = 0; q < 4; ++q) blk[q] -= dq[4 * (size_t)j + q];
  }
}

// Convergence partials: max|r_w/pv|, max|r_o/pv|, sum r_w, sum r_o,
// sum 1/b_w(p), sum 1/b_o(p), sum p, sum pv; with cur_pv the pore volume is
// the current one, pv (1 + c_r (p - p_ref)) (the generic model's criterion).
constexpr int kConvBlocks = 512;
constexpr int kConvQ = 8;
__global__ __launch_bounds__(256) void k_ow_conv(int nc, const double* __restrict__ r, const double* __restrict__ p,
                                                 const double* __restrict__ pv, OWFluid f, int cur_pv,
                                                 double* __restrict__ partial) {
  __shared__ double sh[kConvQ][4];
  const size_t chunk = ((size_t)nc + gridDim.x - 1) / gridDim.x;
  const size_t i0 = (size_t)blockIdx.x * chunk, i1 = min((size_t)nc, i0 + chunk);
  double v[kConvQ] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double rw = r[2 * i], ro = r[2 * i + 1];
    const double q = (cur_pv && f.has_cr) ? pv[i] * (1.0 + f.c_r * (p[i] - f.p_ref)) : pv[i];
    v[7] += q;
    v[0] = fmax(v[0], fabs(rw / q));
    v[1] = fmax(v[1], fabs(ro / q));
    v[2] += rw;
    v[3] += ro;
    double b, db;
    shrink(f.c_w, p[i], f.p_ref, b, db);
    v[4] += 1.0 / b;
    shrink(f.c_o, p[i], f.p_ref, b, db);
    v[5] += 1.0 / b;
    v[6] += p[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    v[0] = fmax(v[0], __shfl_xor(v[0], off, 64));
    v[1] = fmax(v[1], __shfl_xor(v[1], off, 64));
    for (int q = 2; q < kConvQ; ++q) v[q] += __shfl_xor(v[q], off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < kConvQ; ++q) sh[q][wid] = v[q];
  __syncthreads();
  if (threadIdx.x < kConvQ) {
    const int q = threadIdx.x;
    double a = sh[q][0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) a = q < 2 ? fmax(a, sh[q][k]) : a + sh[q][k];
    partial[(size_t)blockIdx.x * kConvQ + q] = a;
  }
}

// One wave per quantity: lanes stride over the block partials, then a
// shuffle tree (fixed order, deterministic); the first 2 quantities are maxima.
__global__ void k_ow_conv_final(int nblk, const double* __restrict__ partial, double* __restrict__ out) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (q >= kConvQ) return;
  const bool is_max = q < 2;
  double a = 0.0;   // the maxima are of absolute values
  for (int b = lane; b < nblk; b += 64) {
    const double v = partial[(size_t)b * kConvQ + q];
    a = is_max ? fmax(a, v) : a + v;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(a, off, 64);
    a = is_max ? fmax(a, o) : a + o;
  }
  if (lane == 0) out[q] = a;
}

// Member-wise convergence partials of a union model: one workgroup per
// chunk of at most kMemChunk cells of one member (a chunk never holds two
// members' cells), the seven quantities of k_ow_conv other than the
// pore-volume sum.  Every partial row is written by each launch.
constexpr int kMemQ = 7;
constexpr int kMemChunk = 2048;
__global__ __launch_bounds__(256) void k_ow_conv_chunks(const int* __restrict__ chunk, const double* __restrict__ r,
                                                        const double* __restrict__ p, const double* __restrict__ pv,
                                                        OWFluid f, double* __restrict__ partial) {
  __shared__ double sh[kMemQ][4];
  const int i0 = chunk[2 * blockIdx.x], i1 = chunk[2 * blockIdx.x + 1];
  double v[kMemQ] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = i0 + (int)threadIdx.x; i < i1; i += (int)blockDim.x) {
    const double rw = r[2 * (size_t)i], ro = r[2 * (size_t)i + 1], q = pv[i];
    v[0] = fmax(v[0], fabs(rw / q));
    v[1] = fmax(v[1], fabs(ro / q));
    v[2] += rw;
    v[3] += ro;
    double b, db;
    shrink(f.c_w, p[i], f.p_ref, b, db);
    v[4] += 1.0 / b;
    shrink(f.c_o, p[i], f.p_ref, b, db);
    v[5] += 1.0 / b;
    v[6] += p[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    v[0] = fmax(v[0], __shfl_xor(v[0], off, 64));
    v[1] = fmax(v[1], __shfl_xor(v[1], off, 64));
    for (int q = 2; q < kMemQ; ++q) v[q] += __shfl_xor(v[q], off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < kMemQ; ++q) sh[q][wid] = v[q];
  __syncthreads();
  if (threadIdx.x < kMemQ) {
    const int q = threadIdx.x;
    double a = sh[q][0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) a = q < 2 ? fmax(a, sh[q][k]) : a + sh[q][k];
    partial[(size_t)blockIdx.x * kMemQ + q] = a;
  }
}

// One wave per member: lanes stride over the member's chunk partials, then a
// fixed shuffle tree per quantity (deterministic).  A member without cells
// gets zeros.
__global__ __launch_bounds__(64) void k_ow_conv_members(const int* __restrict__ ptr,
                                                        const double* __restrict__ partial,
                                                        double* __restrict__ out) {
  const int m = blockIdx.x, lane = threadIdx.x;
  const int c0 = ptr[m], c1 = ptr[m + 1];
  for (int q = 0; q < kMemQ; ++q) {
    const bool is_max = q < 2;
    double a = 0.0;
    for (int c = c0 + lane; c < c1; c += 64) {
      const double v = partial[(size_t)c * kMemQ + q];
      a = is_max ? fmax(a, v) : a + v;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(a, off, 64);
      a = is_max ? fmax(a, o) : a + o;
    }
    if (lane == 0) out[(size_t)m * kMemQ + q] = a;
  }
}

// Newton update with the reference's limits (step :629-651): dp clipped to
// +-dp_max_rel |p|, dsW to +-ds_max, sW to [0, 1]; dx holds sign * update.
__global__ void k_ow_update(int nc, int nw, const double* __restrict__ dx, double sign, double ds_max,
                            double dp_max_rel, double* __restrict__ p, double* __restrict__ s,
                            double* __restrict__ qws, double* __restrict__ qos, double* __restrict__ bhp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc) {
    const double pi = p[i];
    const double cap = dp_max_rel * fabs(pi);
    double dp = sign * dx[2 * (size_t)i];
    double ds = sign * dx[2 * (size_t)i + 1];
    dp = fmin(fmax(dp, -cap), cap);
    ds = fmin(fmax(ds, -ds_max), ds_max);
    p[i] = pi + dp;
    s[i] = fmin(fmax(s[i] + ds, 0.0), 1.0);
  } else if (i < nc + nw) {
    const int w = i - nc;
    const double* d = dx + 2 * (size_t)nc + 3 * (size_t)w;
    qws[w] += sign * d[0];
    qos[w] += sign * d[1];
    bhp[w] += sign * d[2];
  }
}

__global__ void k_negate(size_t n, const double* __restrict__ x, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = -x[i];
}

// ---------------------------------------------------------------------------
OWModelDev::OWModelDev(hipStream_t s) : stream(s) {
  S.B = ST.B = 2;
  S.M = ST.M = 3;
  S.X = ST.X = 2;
  cpr = cpr_create();
}
OWModelDev::~OWModelDev() {
  cpr_destroy(cpr);
  if (cprT) cpr_destroy(cprT);
}

void OWModelDev::set_topology(int nc_, int nconn_, const int* N, const double* T, const double* gdz,
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
  host_bconn = econn;
  // transposed-block map (the pattern is symmetric) and the transposed system
  std::vector<int> tm(nnz);
  for (int i = 0; i < nc; ++i)
    for (int k = pat.rowptr[i]; k < pat.rowptr[i + 1]; ++k) {
      const int j = pat.col[k];
      auto b0 = pat.col.begin() + pat.rowptr[j], b1 = pat.col.begin() + pat.rowptr[j + 1];
      auto it = std::lower_bound(b0, b1, i);
      PMX_CHECK(it != b1 && *it == i, "TPFA pattern is not symmetric");
      tm[k] = (int)(it - pat.col.begin());
    }
  tmap.upload(tm, stream);
  ST.set_pattern(pat, stream);
  if (cprT) {
    cpr_destroy(cprT);
    cprT = nullptr;
  }
  bT.upload(hT, stream);
  bG.upload(hG, stream);
  pv.upload(pv_host, nc, stream);
  w_cpr.alloc(2);
  conv_partial.alloc((size_t)kConvBlocks * kConvQ + 8);
  hier_weights = hT;   // static pressure-graph strength: transmissibility
  PMX_HIP(hipStreamSynchronize(stream));
}

void OWModelDev::set_wells(int nw_, const int* perf_ptr, const int* perf_cells, const double* WI_h,
                           const int* types, const double* vals, const double* compi_h) {
  nw = nw_;
  std::vector<int> ptr(perf_ptr, perf_ptr + nw + 1);
  const int np = nw ? ptr[nw] : 0;
  std::vector<int> cells(perf_cells, perf_cells + np);
  for (int c : cells) PMX_CHECK(c >= 0 && c < nc, "well connection cell out of range");
  S.set_wells(nw, ptr, cells, stream);
  ST.set_wells(nw, ptr, cells, stream);
  WI.upload(WI_h, np, stream);
  cdp.alloc(std::max(np, 1));
  cdp.zero(stream);
  // the two legacy controls: a bhp target (0) or a total surface rate (1)
  std::vector<double> coef(3 * (size_t)nw);
  for (int w = 0; w < nw; ++w) {
    coef[3 * w] = coef[3 * w + 1] = types[w] == 0 ? 0.0 : 1.0;
    coef[3 * w + 2] = types[w] == 0 ? 1.0 : 0.0;
  }
  wctrl.upload(coef.data(), 3 * (size_t)nw, stream);
  wstat.upload(std::vector<int>(nw, 1).data(), nw, stream);
  rscale[0] = rscale[1] = 1.0;
  wval.upload(vals, nw, stream);
  compi.upload(compi_h, 2 * (size_t)nw, stream);
  rq.alloc(2 * (size_t)std::max(np, 1));
  qvol.alloc(2 * (size_t)std::max(np, 1));
  dq.alloc(4 * (size_t)std::max(np, 1));
  mixflag.alloc(std::max(nw, 1));
  PMX_HIP(hipStreamSynchronize(stream));
}

void OWModelDev::set_controls(const double* coef, const double* vals, const int* status, const double* rs) {
  wctrl.upload(coef, 3 * (size_t)nw, stream);
  wval.upload(vals, nw, stream);
  wstat.upload(status, nw, stream);
  rscale[0] = rs[0];
  rscale[1] = rs[1];
  PMX_HIP(hipStreamSynchronize(stream));
}

void OWModelDev::assemble(const double* p, const double* s, const double* qws, const double* qos,
                          const double* bhp, const double* p0, const double* s0, double dt, const double* cdp_dev) {
  PMX_CHECK(nc > 0, "topology not set");
  r.alloc(S.n());
  constexpr int LPR = 8;
  hipLaunchKernelGGL((k_ow_cells<LPR>), dim3(grid_for((size_t)nc * LPR)), dim3(256), 0, stream, nc, S.rowptr.p,
                     S.col.p, S.diag.p, bconn.p, bT.p, bG.p, p, s, p0, s0, pv.p, fluid, 1.0 / dt, dt_cell, r.p, S.val.p);
  PMX_HIP(hipGetLastError());
  if (nw > 0) {
    const double* c = cdp_dev ? cdp_dev : cdp.p;
    hipLaunchKernelGGL(k_ow_wells, dim3(nw), dim3(64), 0, stream, nw, S.well_ptr.p, S.perf_cell.p, WI.p, c, wctrl.p,
                       wstat.p, wval.p, compi.p, p, s, qws, qos, bhp, fluid, rscale[0], rscale[1], (size_t)nc * 2, r.p,
                       rq.p, dq.p, S.Bp.p, S.Cp.p, S.Um.p, S.V.p, S.Dp.p, mixflag.p, qvol.p);
    PMX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ow_perf_scatter, dim3(grid_for(S.npc)), dim3(256), 0, stream, S.npc, S.pc_cell.p,
                       S.pc_ptr.p, S.pc_perf.p, S.diag.p, rq.p, dq.p, r.p, S.val.p);
    PMX_HIP(hipGetLastError());
    // the mixture coupling is active if any well has injecting connections
    // and a live mixture; read back the per-well flags (tiny)
    std::vector<int> mf(nw);
    PMX_HIP(hipMemcpyAsync(mf.data(), mixflag.p, sizeof(int) * nw, hipMemcpyDeviceToHost, stream));
    PMX_HIP(hipStreamSynchronize(stream));
    S.has_mix = false;
    for (int v : mf) S.has_mix |= (v != 0);
  }
}

void OWModelDev::convergence(const double* p, double* out7, double* wells_out) {
  reduce_convergence(p, out7, 7, wells_out, conv_partial.p, kConvBlocks, 0);
}

void OWModelDev::convergence_current_pv(const double* p, double* out8, double* wells_out) {
  reduce_convergence(p, out8, 8, wells_out, conv_partial.p, kConvBlocks, 1);
}

// The reductions into `partial` (nblk rows of 7, then the 7 results) and
// their copy to the host; every partial row and result is written by each
// call, so the buffer needs no clearing between calls.
void OWModelDev::reduce_convergence(const double* p, double* out, int nout, double* wells_out, double* partial,
                                    int nblk, int cur_pv) {
  hipLaunchKernelGGL(k_ow_conv, dim3(nblk), dim3(256), 0, stream, nc, r.p, p, pv.p, fluid, cur_pv, partial);
  PMX_HIP(hipGetLastError());
  double* fin = partial + (size_t)nblk * kConvQ;
  hipLaunchKernelGGL(k_ow_conv_final, dim3(1), dim3(64 * kConvQ), 0, stream, nblk, partial, fin);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipMemcpyAsync(out, fin, sizeof(double) * nout, hipMemcpyDeviceToHost, stream));
  if (nw > 0 && wells_out)
    PMX_HIP(hipMemcpyAsync(wells_out, r.p + 2 * (size_t)nc, sizeof(double) * 3 * nw, hipMemcpyDeviceToHost, stream));
  PMX_HIP(hipStreamSynchronize(stream));
}

// Allocation-free convergence reductions.  The partials live in `scratch`,
// one copy per GPU for the life of the process: kConvBlocks rows of 8
// doubles plus the 8 results (32,832 bytes).  Its device address is looked
// up once per GPU, with that GPU current, and only that address is handed
// to kernels and copies.  Calls that overlap on one GPU share the array and
// race.
constexpr size_t kConvScratchDoubles = (size_t)kConvBlocks * kConvQ + 8;
static __device__ __align__(256) double scratch[kConvScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static double* scratch_address(int device) {
  PMX_CHECK(device >= 0 && device < kMaxDevices, "device id out of range for the static scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    void* ptr = nullptr;
    PMX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    PMX_CHECK(ptr != nullptr, "no device address for the static scratch array");
    g_scratch_ptr[device] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[device];
}

void OWModelDev::convergence_noalloc(int device, const double* p, double* out7, double* wells_out) {
  PMX_CHECK(nc > 0 && r.n >= (size_t)2 * nc, "assemble before the convergence reductions");
  // one partial row per 2,048 cells, at most kConvBlocks rows
  const int nblk = (int)std::min<size_t>(kConvBlocks, ((size_t)nc + 2047) / 2048);
  PMX_CHECK((size_t)nblk * kConvQ + kConvQ <= kConvScratchDoubles,
            "convergence reductions need more partial rows than the static scratch holds");
  reduce_convergence(p, out7, 7, wells_out, scratch_address(device), nblk, 0);
}

int OWModelDev::solve(const LinearSolverOptions& o, double* dx, SolveInfo& info, double pmean) {
  auto t0 = std::chrono::steady_clock::now();
  if (!cpr_has_hierarchy(cpr)) cpr_build_hierarchy(cpr, S, hier_weights, o, stream);
  // quasi-IMPES weights 1/b_i at the mean pressure (linear_solvers.py:155-190)
  double w[2] = {1.0, 1.0};
  if (fluid.c_w != 0.0) w[0] = 1.0 / std::exp(fluid.c_w * (pmean - fluid.p_ref));
  if (fluid.c_o != 0.0) w[1] = 1.0 / std::exp(fluid.c_o * (pmean - fluid.p_ref));
  PMX_HIP(hipMemcpyAsync(w_cpr.p, w, sizeof(w), hipMemcpyHostToDevice, stream));
  cpr_setup(cpr, S, w_cpr.p, o, stream);
  PMX_HIP(hipStreamSynchronize(stream));
  const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rhs.alloc(S.n());
  hipLaunchKernelGGL(k_negate, dim3(grid_for(S.n())), dim3(256), 0, stream, S.n(), r.p, rhs.p);
  PMX_HIP(hipGetLastError());
  const int st = cpr_solve(cpr, S, rhs.p, dx, o, info, stream);
  info.setup_ms = setup_ms;
  return st;
}

__global__ void k_mask(size_t n, const double* __restrict__ mask, double* __restrict__ r) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] *= mask[i];
}

void OWModelDev::set_members(int nm, const int* cell_ptr) {
  PMX_CHECK(nm >= 1 && cell_ptr[0] == 0 && cell_ptr[nm] == nc, "member cell ranges must cover the cells");
  std::vector<int> chunk, ptr(1, 0);
  for (int m = 0; m < nm; ++m) {
    PMX_CHECK(cell_ptr[m + 1] >= cell_ptr[m], "member cell ranges must be ascending");
    for (int i0 = cell_ptr[m]; i0 < cell_ptr[m + 1]; i0 += kMemChunk) {
      chunk.push_back(i0);
      chunk.push_back(std::min(i0 + kMemChunk, cell_ptr[m + 1]));
    }
    ptr.push_back((int)chunk.size() / 2);
  }
  nmem = nm;
  nchunk = ptr.back();
  mem_chunk.upload(chunk, stream);
  mem_chunk_ptr.upload(ptr, stream);
  mem_partial.alloc((size_t)std::max(nchunk, 1) * kMemQ);
  mem_out.alloc((size_t)nmem * kMemQ);
  PMX_HIP(hipStreamSynchronize(stream));   // the host tables above go out of scope
}

void OWModelDev::member_convergence(const double* p, double* out7, double* wells_out) {
  PMX_CHECK(nmem > 0, "set the members before their convergence reductions");
  PMX_CHECK(r.n >= S.n(), "assemble before the convergence reductions");
  if (nchunk > 0) {
    hipLaunchKernelGGL(k_ow_conv_chunks, dim3(nchunk), dim3(256), 0, stream, mem_chunk.p, r.p, p, pv.p, fluid,
                       mem_partial.p);
    PMX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_ow_conv_members, dim3(nmem), dim3(64), 0, stream, mem_chunk_ptr.p, mem_partial.p,
                     mem_out.p);
  PMX_HIP(hipGetLastError());
  PMX_HIP(hipMemcpyAsync(out7, mem_out.p, sizeof(double) * kMemQ * nmem, hipMemcpyDeviceToHost, stream));
  if (nw > 0 && wells_out)
    PMX_HIP(hipMemcpyAsync(wells_out, r.p + 2 * (size_t)nc, sizeof(double) * 3 * nw, hipMemcpyDeviceToHost, stream));
  PMX_HIP(hipStreamSynchronize(stream));
}

void OWModelDev::mask_residual(const double* mask) {
  PMX_CHECK(r.n >= S.n(), "assemble before masking the residual");
  hipLaunchKernelGGL(k_mask, dim3(grid_for(S.n())), dim3(256), 0, stream, S.n(), mask, r.p);
  PMX_HIP(hipGetLastError());
}

void OWModelDev::update(const double* dx, double sign, double ds_max, double dp_max_rel, double* p, double* s,
                        double* qws, double* qos, double* bhp) {
  hipLaunchKernelGGL(k_ow_update, dim3(grid_for((size_t)nc + nw)), dim3(256), 0, stream, nc, nw, dx, sign, ds_max,
                     dp_max_rel, p, s, qws, qos, bhp);
  PMX_HIP(hipGetLastError());
}

}  // namespace pmx
