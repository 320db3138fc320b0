// Discrete adjoint of the oil-water Newton step on the device.
//
// Reference: pymrst/modules/optimization/adjoint.py:73-225 (backward sweep
// (dR_n/dx_n)^T lam_n = -dJ_n/dx_n - (dR_{n+1}/dx_n)^T lam_{n+1}, with
// dR_n/dx_{n-1} from the accumulation term only, :85-107).  The reference
// differentiates with respect to well controls; this file adds the
// log-permeability derivative of the residual: transmissibilities
// T_f = 1/(1/T_h0 + 1/T_h1) with T_h proportional to k (dT_f/dln k_c =
// T_f^2/T_h,c) and Peaceman well indices proportional to k.
//
// The transposed system keeps the forward BCSR pattern (TPFA patterns are
// symmetric): block (i,j) of J^T is block (j,i) of J, transposed, and the
// well parts swap roles (Bp <-> Cp^T, Um <-> V^T, D -> D^T).  It is solved
// with the same CPR-FGMRES, restricting with e_0 (the pressure-variable rows
// of J^T) and prolonging with the quasi-IMPES weights, so that its pressure
// matrix is (W J P)^T.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>

#include "ow.h"
#include "ow_props.h"

namespace pmx {

template <int B>
__global__ void k_transpose_blocks(int nnzb, const int* __restrict__ tmap, const double* __restrict__ val,
                                   double* __restrict__ valT) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnzb) return;
  const size_t q = (size_t)tmap[k];
#pragma unroll
  for (int a = 0; a < B; ++a)
#pragma unroll
    for (int b = 0; b < B; ++b) valT[q * B * B + b * B + a] = val[(size_t)k * B * B + a * B + b];
}

// Per perforation: Bp^T = Cp^T (BxM <- MxB), Cp^T = Bp^T, Um^T = V^T (BxX <-
// XxB), V^T = Um^T; per well: D^T.
template <int B, int M, int X>
__global__ void k_transpose_wells(int nperf, int nw, const double* __restrict__ Bp, const double* __restrict__ Cp,
                                  const double* __restrict__ Um, const double* __restrict__ V,
                                  const double* __restrict__ Dp, double* __restrict__ BpT, double* __restrict__ CpT,
                                  double* __restrict__ UmT, double* __restrict__ VT, double* __restrict__ DpT) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nperf) {
    for (int a = 0; a < B; ++a)
      for (int q = 0; q < M; ++q) {
        BpT[(size_t)j * B * M + a * M + q] = Cp[(size_t)j * M * B + q * B + a];
        CpT[(size_t)j * M * B + q * B + a] = Bp[(size_t)j * B * M + a * M + q];
      }
    for (int a = 0; a < B; ++a)
      for (int x = 0; x < X; ++x) {
        UmT[(size_t)j * B * X + a * X + x] = V[(size_t)j * X * B + x * B + a];
        VT[(size_t)j * X * B + x * B + a] = Um[(size_t)j * B * X + a * X + x];
      }
  }
  if (j < nw)
    for (int p = 0; p < M; ++p)
      for (int q = 0; q < M; ++q) DpT[(size_t)j * M * M + q * M + p] = Dp[(size_t)j * M * M + p * M + q];
}

// out = (dR/dx_prev)^T lam: the accumulation term's derivative with respect
// to (p0, sW0), a 2x2 block per cell; zero on the well unknowns.
__global__ void k_ow_prev_transpose(int nc, int nw, const double* __restrict__ p0, const double* __restrict__ s0,
                                    const double* __restrict__ pv, OWFluid f, double inv_dt,
                                    const double* __restrict__ lam, double* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) {
    const double pp0 = p0[c], ss0 = s0[c], v = pv[c];
    double bw0, dbw0, bo0, dbo0;
    shrink(f.c_w, pp0, f.p_ref, bw0, dbw0);
    shrink(f.c_o, pp0, f.p_ref, bo0, dbo0);
    double m0 = 1.0, dm = 0.0;
    if (f.has_cr) {
      m0 = 1.0 + f.c_r * (pp0 - f.p_ref);
      dm = f.c_r;
    }
    const double pvm0 = v * m0;
    const double a00 = -v * (dm * bw0 + m0 * dbw0) * ss0 * inv_dt;   // d water / d p0
    const double a01 = -pvm0 * bw0 * inv_dt;                          // d water / d sW0
    const double a10 = -v * (dm * bo0 + m0 * dbo0) * (1.0 - ss0) * inv_dt;
    const double a11 = pvm0 * bo0 * inv_dt;
    const double lw = lam[2 * (size_t)c], lo = lam[2 * (size_t)c + 1];
    out[2 * (size_t)c] = a00 * lw + a10 * lo;
    out[2 * (size_t)c + 1] = a01 * lw + a11 * lo;
  } else if (c < nc + 3 * nw) {
    out[2 * (size_t)nc + (c - nc)] = 0.0;
  }
}

// Flux part of (dR/dln k)^T lam, cell-parallel (LPR lanes per row): each
// connection's phase fluxes C_a (as seen from the row) are linear in T_f, so
//   grad_c += sum_f sum_a (lam_a,c - lam_a,n) C_a,f * T_f / T_h,c.
template <int LPR>
__global__ __launch_bounds__(256) void k_ow_dlogk_cells(int nc, const int* __restrict__ rowptr,
                                                        const int* __restrict__ col, const int* __restrict__ bconn,
                                                        const double* __restrict__ bT, const double* __restrict__ bG,
                                                        const double* __restrict__ bfrac,
                                                        const double* __restrict__ p, const double* __restrict__ s,
                                                        OWFluid f, const double* __restrict__ lam,
                                                        double* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  double g = 0.0;
  if (row < nc) {
    const double pr = p[row], sr = s[row];
    const CellProps P = cell_props(f, pr, sr);
    const double lwr = lam[2 * (size_t)row], lor = lam[2 * (size_t)row + 1];
    for (int k = rowptr[row] + lane; k < rowptr[row + 1]; k += LPR) {
      const int c = col[k];
      if (c == row) continue;
      const double pn = p[c], sn = s[c];
      const CellProps Q = cell_props(f, pn, sn);
      const bool isN0 = (bconn[k] & 1) == 0;
      double Cw, Co, d0, d1, d2, d3;
      phase_flux(bT[k], bG[k], isN0, pr, pn, f.rho_w, P.bw, P.dbw, P.mw, P.dmw, Q.bw, Q.dbw, Q.mw, Q.dmw, Cw, d0, d1,
                 d2, d3);
      phase_flux(bT[k], bG[k], isN0, pr, pn, f.rho_o, P.bo, P.dbo, P.mo, P.dmo, Q.bo, Q.dbo, Q.mo, Q.dmo, Co, d0, d1,
                 d2, d3);
      g += ((lwr - lam[2 * (size_t)c]) * Cw + (lor - lam[2 * (size_t)c + 1]) * Co) * bfrac[k];
    }
  }
  g = lane_sum(g, LPR);
  if (row < nc && lane == 0) grad[row] += g;
}

// Well part of (dR/dln k)^T lam, one 64-wide block per well.  Connection
// rates are linear in their own WI; through the wellbore mixture, the rates
// of injecting connections also depend on the WI of producing ones:
//   g_j = -mu_j . q_j                       (mu_j = lam_cell(c_j) + lam_well[0:2])
//   g_k += nu^T A q_k  for producing k       (nu = sum_{j injecting} U_j^T mu_j,
//                                             A = d mix / d wbq)
// Output per connection; the cell scatter follows.
__global__ __launch_bounds__(64) void k_ow_dlogk_wells(
    int nw, const int* __restrict__ well_ptr, const int* __restrict__ perf_cell, const double* __restrict__ WI,
    const double* __restrict__ cdp, const double* __restrict__ compi, const double* __restrict__ p,
    const double* __restrict__ s, const double* __restrict__ qws, const double* __restrict__ qos,
    const double* __restrict__ bhp, OWFluid f, int nc, const double* __restrict__ lam, double* __restrict__ gperf) {
  const int w = blockIdx.x, lane = threadIdx.x;
  const int j0 = well_ptr[w], j1 = well_ptr[w + 1];
  const double bh = bhp[w], qw = qws[w], qo = qos[w];
  const double lww = lam[2 * (size_t)nc + 3 * (size_t)w], lwo = lam[2 * (size_t)nc + 3 * (size_t)w + 1];
  double Sw = 0, So = 0;
  int anyprod = 0;
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const double dd = -(bh + cdp[j]) + p[c];
    if (dd < 0.0) continue;
    anyprod = 1;
    const CellProps P = cell_props(f, p[c], s[c]);
    Sw += P.bw * (-WI[j] * P.mw * dd);
    So += P.bo * (-WI[j] * P.mo * dd);
  }
  Sw = lane_sum(Sw, 64);
  So = lane_sum(So, 64);
  anyprod = __any(anyprod) ? 1 : 0;
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
  double nu0 = 0.0, nu1 = 0.0;
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const double pc = p[c];
    const double dd = -(bh + cdp[j]) + pc;
    const CellProps P = cell_props(f, pc, s[c]);
    const double t = WI[j];
    const double mu0 = lam[2 * (size_t)c] + lww, mu1 = lam[2 * (size_t)c + 1] + lwo;
    double cw, co;
    if (dd < 0.0) {
      const double mt = P.mw + P.mo;
      const double cqt = -t * (mt * dd);
      const double vr = mixw / P.bw + mixo / P.bo;
      const double cis = cqt / vr;
      const double q2 = cqt / (vr * vr);
      cw = mixw * cis;
      co = mixo * cis;
      const double U00 = cis - mixw * q2 / P.bw, U01 = -mixw * q2 / P.bo;
      const double U10 = -mixo * q2 / P.bw, U11 = cis - mixo * q2 / P.bo;
      nu0 += U00 * mu0 + U10 * mu1;
      nu1 += U01 * mu0 + U11 * mu1;
    } else {
      cw = P.bw * (-t * P.mw * dd);
      co = P.bo * (-t * P.mo * dd);
    }
    gperf[j] = -(mu0 * cw + mu1 * co);
  }
  nu0 = lane_sum(nu0, 64);
  nu1 = lane_sum(nu1, 64);
  if (dead) return;
  // A q_k for producing connections, contracted with nu
  const double a0 = nu0 * aww + nu1 * aow, a1 = nu0 * awo + nu1 * aoo;   // nu^T A
  for (int j = j0 + lane; j < j1; j += 64) {
    const int c = perf_cell[j];
    const double pc = p[c];
    const double dd = -(bh + cdp[j]) + pc;
    if (dd < 0.0) continue;
    const CellProps P = cell_props(f, pc, s[c]);
    const double t = WI[j];
    const double cw = P.bw * (-t * P.mw * dd), co = P.bo * (-t * P.mo * dd);
    gperf[j] += a0 * cw + a1 * co;
  }
}

__global__ void k_perf_gather_add(int npc, const int* __restrict__ pc_cell, const int* __restrict__ pc_ptr,
                                  const int* __restrict__ pc_perf, const double* __restrict__ gperf,
                                  double* __restrict__ grad) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npc) return;
  double g = 0.0;
  for (int t = pc_ptr[u]; t < pc_ptr[u + 1]; ++t) g += gperf[pc_perf[t]];
  grad[pc_cell[u]] += g;
}

// ---------------------------------------------------------------------------
void OWModelDev::set_trans_fractions(const double* frac) {
  PMX_CHECK(nc > 0, "topology not set");
  const int nnz = S.nnzb;
  std::vector<double> h(nnz, 0.0);
  for (int k = 0; k < nnz; ++k) {
    const int code = host_bconn[k];
    if (code < 0) continue;
    h[k] = frac[2 * (code >> 1) + (code & 1)];
  }
  bfrac.upload(h, stream);
  PMX_HIP(hipStreamSynchronize(stream));
}

void OWModelDev::transpose_system() {
  ST.has_mix = S.has_mix;
  hipLaunchKernelGGL((k_transpose_blocks<2>), dim3(grid_for(S.nnzb)), dim3(256), 0, stream, S.nnzb, tmap.p, S.val.p,
                     ST.val.p);
  PMX_HIP(hipGetLastError());
  if (nw > 0) {
    const int np = S.nperf;
    hipLaunchKernelGGL((k_transpose_wells<2, 3, 2>), dim3(grid_for(std::max(np, nw))), dim3(256), 0, stream, np, nw,
                       S.Bp.p, S.Cp.p, S.Um.p, S.V.p, S.Dp.p, ST.Bp.p, ST.Cp.p, ST.Um.p, ST.V.p, ST.Dp.p);
    PMX_HIP(hipGetLastError());
  }
}

int OWModelDev::adjoint_solve(const LinearSolverOptions& o, double pmean, const double* rhs_dev, double* lam,
                              SolveInfo& info) {
  auto t0 = std::chrono::steady_clock::now();
  transpose_system();
  if (cprT == nullptr) cprT = cpr_create();
  if (!cpr_has_hierarchy(cprT)) {
    if (!line_ptr.empty()) cpr_set_lines(cprT, (int)line_ptr.size() - 1, line_ptr.data(), line_cells.data());
    cpr_build_hierarchy(cprT, ST, hier_weights, o, stream);
  }
  // restrict with the pressure-variable rows, prolong with the weights
  double w[2] = {1.0, 0.0}, pc[2] = {1.0, 1.0};
  if (fluid.c_w != 0.0) pc[0] = 1.0 / std::exp(fluid.c_w * (pmean - fluid.p_ref));
  if (fluid.c_o != 0.0) pc[1] = 1.0 / std::exp(fluid.c_o * (pmean - fluid.p_ref));
  double h[4] = {w[0], w[1], pc[0], pc[1]};
  w_adj.alloc(4);
  PMX_HIP(hipMemcpyAsync(w_adj.p, h, sizeof(h), hipMemcpyHostToDevice, stream));
  cpr_setup(cprT, ST, w_adj.p, w_adj.p + 2, o, stream);
  PMX_HIP(hipStreamSynchronize(stream));
  const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const int st = cpr_solve(cprT, ST, rhs_dev, lam, o, info, stream);
  info.setup_ms = setup_ms;
  return st;
}

void OWModelDev::prev_state_transpose(const double* p0, const double* s0, double dt, const double* lam,
                                      double* out) {
  hipLaunchKernelGGL(k_ow_prev_transpose, dim3(grid_for((size_t)nc + 3 * nw)), dim3(256), 0, stream, nc, nw, p0, s0,
                     pv.p, fluid, 1.0 / dt, lam, out);
  PMX_HIP(hipGetLastError());
}

void OWModelDev::dlogk(const double* p, const double* s, const double* qws, const double* qos, const double* bhp,
                       const double* cdp_dev, const double* lam, double* grad) {
  PMX_CHECK(bfrac.n == (size_t)S.nnzb, "transmissibility fractions not set");
  constexpr int LPR = 8;
  hipLaunchKernelGGL((k_ow_dlogk_cells<LPR>), dim3(grid_for((size_t)nc * LPR)), dim3(256), 0, stream, nc,
                     S.rowptr.p, S.col.p, bconn.p, bT.p, bG.p, bfrac.p, p, s, fluid, lam, grad);
  PMX_HIP(hipGetLastError());
  if (nw > 0) {
    gperf.alloc(std::max(S.nperf, 1));
    const double* c = cdp_dev ? cdp_dev : cdp.p;
    hipLaunchKernelGGL(k_ow_dlogk_wells, dim3(nw), dim3(64), 0, stream, nw, S.well_ptr.p, S.perf_cell.p, WI.p, c,
                       compi.p, p, s, qws, qos, bhp, fluid, nc, lam, gperf.p);
    PMX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_perf_gather_add, dim3(grid_for(S.npc)), dim3(256), 0, stream, S.npc, S.pc_cell.p,
                       S.pc_ptr.p, S.pc_perf.p, gperf.p, grad);
    PMX_HIP(hipGetLastError());
  }
}

}  // namespace pmx
