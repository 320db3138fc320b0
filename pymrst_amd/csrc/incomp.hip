// Incompressible TPFA pressure solve on the device, fed by a computeTrans
// kernel.
//
// computeTrans: pymrst/solvers/compute_trans.py:11-29 -- per half-face
//   T = sum_ij c_i K_ij n_j / |c|^2, c = face centroid - cell centroid,
//   n = area normal pointing out of the cell, nine tensor terms summed in the
//   reference's order; negative values replaced by their magnitude (:32-42).
// incompTPFA: incomp_tpfa_ad, pymrst/modules/incomp/incomp_tpfa.py:252-389 --
//   mobility-weighted half-face T, harmonic face T, gravity / Dirichlet /
//   Neumann / source right-hand side (:392-435), bhp and rate wells as extra
//   unknowns, null-space pin without Dirichlet or bhp wells, face pressure and
//   flux reconstruction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>

#include "incomp.h"

namespace pmx {

__global__ void k_compute_trans(int nhf, const int* __restrict__ hf_cell, const int* __restrict__ hf_face,
                                const int* __restrict__ face_nb, const double* __restrict__ cc,
                                const double* __restrict__ fc, const double* __restrict__ fn, int nk,
                                const double* __restrict__ perm, double* __restrict__ T) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nhf) return;
  const int c = hf_cell[h], f = hf_face[h];
  double cv[3], nv[3];
  const double sgn = face_nb[2 * f] == c ? 1.0 : -1.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cv[d] = fc[3 * (size_t)f + d] - cc[3 * (size_t)c + d];
    nv[d] = sgn * fn[3 * (size_t)f + d];
  }
  // symmetric tensor (xx, xy, xz, yy, yz, zz) from 1, 3 or 6 columns
  double t6[6];
  const double* k = perm + (size_t)c * nk;
  if (nk == 1) {
    t6[0] = k[0]; t6[1] = 0.0; t6[2] = 0.0; t6[3] = k[0]; t6[4] = 0.0; t6[5] = k[0];
  } else if (nk == 3) {
    t6[0] = k[0]; t6[1] = 0.0; t6[2] = 0.0; t6[3] = k[1]; t6[4] = 0.0; t6[5] = k[2];
  } else {
#pragma unroll
    for (int q = 0; q < 6; ++q) t6[q] = k[q];
  }
  const int idx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) t += cv[i / 3] * t6[idx[i]] * nv[i % 3];
  t /= (cv[0] * cv[0] + cv[1] * cv[1]) + cv[2] * cv[2];
  T[h] = fabs(t);
}

// Per half-face: mobility-weighted T and gravity potential omega (x_f - x_c).g
__global__ void k_incomp_halfface(int nhf, const int* __restrict__ hf_cell, const int* __restrict__ hf_face,
                                  const double* __restrict__ T, const double* __restrict__ mob,
                                  const double* __restrict__ omega, const double* __restrict__ cc,
                                  const double* __restrict__ fc, double gx, double gy, double gz, int grav_on,
                                  double* __restrict__ thf, double* __restrict__ ghf) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nhf) return;
  const int c = hf_cell[h], f = hf_face[h];
  thf[h] = T[h] * mob[c];
  double g = 0.0;
  if (grav_on) {
    const double cx = fc[3 * (size_t)f] - cc[3 * (size_t)c];
    const double cy = fc[3 * (size_t)f + 1] - cc[3 * (size_t)c + 1];
    const double cz = fc[3 * (size_t)f + 2] - cc[3 * (size_t)c + 2];
    g = omega[c] * (cx * gx + cy * gy + cz * gz);
  }
  ghf[h] = g;
}

// Per face: harmonic T, signed gravity sum fg (internal and Dirichlet faces)
// and the unsigned sum over all half-faces (boundary flux reconstruction).
__global__ void k_incomp_face(int nf, const int* __restrict__ face_hf, const int* __restrict__ face_nb,
                              const int* __restrict__ hf_cell, const double* __restrict__ thf,
                              const double* __restrict__ ghf, const unsigned char* __restrict__ dF,
                              double* __restrict__ ft, double* __restrict__ fg, double* __restrict__ fg_all) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const bool internal = face_nb[2 * f] >= 0 && face_nb[2 * f + 1] >= 0;
  double inv = 0.0, s = 0.0, a = 0.0;
  for (int q = 0; q < 2; ++q) {
    const int h = face_hf[2 * f + q];
    if (h < 0) continue;
    inv += 1.0 / thf[h];
    const double sg = face_nb[2 * f] == hf_cell[h] ? 1.0 : -1.0;
    if (internal || dF[f]) s += ghf[h] * sg;
    a += ghf[h];
  }
  ft[f] = 1.0 / inv;
  fg[f] = s;
  fg_all[f] = a;
}

// Cell rows of the pressure system (one thread per cell over its half-faces),
// values into the BCSR (1x1) pattern and the right-hand side.
__global__ void k_incomp_rows(int nc, const int* __restrict__ face_pos, const int* __restrict__ hf_face,
                              const int* __restrict__ face_nb, const double* __restrict__ thf,
                              const double* __restrict__ ft, const double* __restrict__ fg,
                              const unsigned char* __restrict__ dF, const double* __restrict__ dval,
                              const double* __restrict__ hh, const double* __restrict__ gg,
                              const int* __restrict__ rowptr, const int* __restrict__ col,
                              const int* __restrict__ diag, double* __restrict__ val, double* __restrict__ rhs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double r = 0.0, d = 0.0;
  for (int h = face_pos[c]; h < face_pos[c + 1]; ++h) {
    const int f = hf_face[h];
    const int n0 = face_nb[2 * f], n1 = face_nb[2 * f + 1];
    const double sgn = n0 == c ? 1.0 : -1.0;
    const double ffh = dF[f] ? -dval[f] : 0.0;
    r += -ft[f] * (sgn * fg[f] + ffh);
    r += -hh[f];
    if (n0 >= 0 && n1 >= 0) {
      d += ft[f];
      const int other = n0 == c ? n1 : n0;
      // locate the entry (row c, column other)
      int lo = rowptr[c], hi = rowptr[c + 1] - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < other) lo = mid + 1; else hi = mid;
      }
      val[lo] = -ft[f];
    } else if (dF[f]) {
      d += thf[h];
    }
  }
  val[diag[c]] = d;
  rhs[c] = r + gg[c];
}

// Well connections: add wi to the perforated cells' diagonals and their
// right-hand-side contributions (unique perforated cells, inputs in order).
__global__ void k_incomp_perfs(int npc, const int* __restrict__ pc_cell, const int* __restrict__ pc_ptr,
                               const int* __restrict__ pc_perf, const int* __restrict__ diag,
                               const double* __restrict__ wi, const double* __restrict__ radd,
                               double* __restrict__ val, double* __restrict__ rhs) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npc) return;
  const int c = pc_cell[u];
  for (int t = pc_ptr[u]; t < pc_ptr[u + 1]; ++t) {
    const int j = pc_perf[t];
    val[diag[c]] += wi[j];
    rhs[c] += radd[j];
  }
}

// Face pressures and fluxes (incomp_tpfa.py:365-381).
__global__ void k_incomp_reconstruct(int nf, const int* __restrict__ face_hf, const int* __restrict__ face_nb,
                                     const int* __restrict__ hf_cell, const double* __restrict__ thf,
                                     const double* __restrict__ ghf, const double* __restrict__ ft,
                                     const double* __restrict__ fg, const double* __restrict__ fg_all,
                                     const unsigned char* __restrict__ dF, const double* __restrict__ dval,
                                     const double* __restrict__ hh, const double* __restrict__ p,
                                     double* __restrict__ fpress, double* __restrict__ flux) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double num = 0.0, den = 0.0;
  for (int q = 0; q < 2; ++q) {
    const int h = face_hf[2 * f + q];
    if (h < 0) continue;
    num += (p[hf_cell[h]] + ghf[h]) * thf[h];
    den += thf[h];
  }
  const int n0 = face_nb[2 * f], n1 = face_nb[2 * f + 1];
  double fp = num / den;
  if (n0 < 0 || n1 < 0) fp -= hh[f] / ft[f];
  if (dF[f]) fp = dval[f];
  fpress[f] = fp;
  if (n0 >= 0 && n1 >= 0) {
    flux[f] = -ft[f] * (p[n1] - p[n0] - fg[f]);
  } else {
    const double bs = n1 < 0 ? 1.0 : -1.0;
    const int c = n0 >= 0 ? n0 : n1;
    flux[f] = -bs * ft[f] * (fp - p[c] - fg_all[f]);
  }
}

// Periodic-boundary pressure jumps (incomp_tpfa.py:113-118): added to the
// gravity term of their faces, so they enter the right-hand side and the
// interior fluxes.
__global__ void k_add_face_jump(int n, const int* __restrict__ face, const double* __restrict__ jump,
                                double* __restrict__ fg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fg[face[i]] += jump[i];
}

__global__ void k_scale_entry(double* __restrict__ val, int k, double a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) val[k] *= a;
}

// ---------------------------------------------------------------------------
IncompDev::IncompDev(hipStream_t s) : stream(s) {
  S.B = 1;
  S.M = 1;
  S.X = 1;
  cpr = cpr_create();
}
IncompDev::~IncompDev() { cpr_destroy(cpr); }

void IncompDev::set_grid(int nc_, int nf_, const int* face_pos_h, const int* hf_face_h, const int* face_nb_h,
                         const double* cc_h, const double* fc_h) {
  nc = nc_;
  nf = nf_;
  nhf = face_pos_h[nc];
  std::vector<int> hf_cell_h(nhf), face_hf_h(2 * (size_t)nf, -1);
  for (int c = 0; c < nc; ++c)
    for (int h = face_pos_h[c]; h < face_pos_h[c + 1]; ++h) hf_cell_h[h] = c;
  for (int h = 0; h < nhf; ++h) {
    const int f = hf_face_h[h];
    PMX_CHECK(f >= 0 && f < nf, "half-face refers to a face out of range");
    const int c = hf_cell_h[h];
    const int side = face_nb_h[2 * f] == c ? 0 : 1;
    PMX_CHECK(face_nb_h[2 * f + side] == c, "half-face cell is not a neighbour of its face");
    face_hf_h[2 * (size_t)f + side] = h;
  }
  std::vector<int> N;
  for (int f = 0; f < nf; ++f)
    if (face_nb_h[2 * f] >= 0 && face_nb_h[2 * f + 1] >= 0) {
      N.push_back(face_nb_h[2 * f]);
      N.push_back(face_nb_h[2 * f + 1]);
    }
  HostPattern pat;
  std::vector<int> econn;
  tpfa_pattern(nc, (int)N.size() / 2, N.data(), pat, econn);
  S.set_pattern(pat, stream);
  host_pat = pat;
  face_pos.upload(face_pos_h, nc + 1, stream);
  hf_face.upload(hf_face_h, nhf, stream);
  hf_cell.upload(hf_cell_h, stream);
  face_nb.upload(face_nb_h, 2 * (size_t)nf, stream);
  face_hf.upload(face_hf_h, stream);
  cc.upload(cc_h, 3 * (size_t)nc, stream);
  fc.upload(fc_h, 3 * (size_t)nf, stream);
  thf.alloc(nhf);
  ghf.alloc(nhf);
  ft.alloc(nf);
  fg.alloc(nf);
  fg_all.alloc(nf);
  dF.alloc(nf);
  dval.alloc(nf);
  hh.alloc(nf);
  gg.alloc(nc);
  w1.alloc(1);
  const double one = 1.0;
  PMX_HIP(hipMemcpyAsync(w1.p, &one, sizeof(double), hipMemcpyHostToDevice, stream));
  PMX_HIP(hipStreamSynchronize(stream));
}

void IncompDev::compute_trans(const double* fn, int nk, const double* perm, double* T) {
  hipLaunchKernelGGL(k_compute_trans, dim3(grid_for(nhf)), dim3(256), 0, stream, nhf, hf_cell.p, hf_face.p,
                     face_nb.p, cc.p, fc.p, fn, nk, perm, T);
  PMX_HIP(hipGetLastError());
}

void IncompDev::assemble(const double* T, const double* mob, const double* omega, const double* g,
                         const IncompBC& bc, const IncompWells& W) {
  // boundary conditions and sources (compute_pressure_rhs semantics)
  std::vector<unsigned char> dF_h(nf, 0);
  std::vector<double> dval_h(nf, 0.0), hh_h(nf, 0.0), gg_h(nc, 0.0);
  for (size_t i = 0; i < bc.face.size(); ++i) {
    const int f = bc.face[i];
    PMX_CHECK(f >= 0 && f < nf, "boundary condition on a face out of range");
    if (bc.type[i] == 0) {
      PMX_CHECK(!dF_h[f], "repeated faces in boundary condition");
      dF_h[f] = 1;
      dval_h[f] = bc.value[i];
    } else {
      hh_h[f] = -bc.value[i];
    }
  }
  for (size_t i = 0; i < bc.src_cell.size(); ++i) {
    const int c = bc.src_cell[i];
    PMX_CHECK(c >= 0 && c < nc, "source in a cell out of range");
    gg_h[c] += bc.src_rate[i];
  }
  dF.upload(dF_h, stream);
  dval.upload(dval_h, stream);
  hh.upload(hh_h, stream);
  gg.upload(gg_h, stream);
  const bool grav_on = g && (g[0] != 0.0 || g[1] != 0.0 || g[2] != 0.0);
  hipLaunchKernelGGL(k_incomp_halfface, dim3(grid_for(nhf)), dim3(256), 0, stream, nhf, hf_cell.p, hf_face.p, T,
                     mob, omega, cc.p, fc.p, grav_on ? g[0] : 0.0, grav_on ? g[1] : 0.0, grav_on ? g[2] : 0.0,
                     (int)grav_on, thf.p, ghf.p);
  PMX_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_incomp_face, dim3(grid_for(nf)), dim3(256), 0, stream, nf, face_hf.p, face_nb.p, hf_cell.p,
                     thf.p, ghf.p, dF.p, ft.p, fg.p, fg_all.p);
  PMX_HIP(hipGetLastError());
  if (!bc.jump_face.empty()) {
    for (int f : bc.jump_face) PMX_CHECK(f >= 0 && f < nf, "periodic boundary condition on a face out of range");
    jump_face.upload(bc.jump_face, stream);
    jump_val.upload(bc.jump_value, stream);
    hipLaunchKernelGGL(k_add_face_jump, dim3(grid_for(bc.jump_face.size())), dim3(256), 0, stream,
                       (int)bc.jump_face.size(), jump_face.p, jump_val.p, fg.p);
    PMX_HIP(hipGetLastError());
  }
  // wells: host part (per connection wi = WI mob, right-hand sides)
  const int nw = (int)W.type.size();
  if (nw != S.nw || (nw && W.perf_ptr[nw] != S.nperf)) {
    S.set_wells(nw, W.perf_ptr, W.cells, stream);
    wi_d.alloc(std::max(S.nperf, 1));
    radd_d.alloc(std::max(S.nperf, 1));
  }
  r.alloc(S.n());
  hipLaunchKernelGGL(k_incomp_rows, dim3(grid_for(nc)), dim3(256), 0, stream, nc, face_pos.p, hf_face.p, face_nb.p,
                     thf.p, ft.p, fg.p, dF.p, dval.p, hh.p, gg.p, S.rowptr.p, S.col.p, S.diag.p, S.val.p, r.p);
  PMX_HIP(hipGetLastError());
  std::vector<double>& wi = wi_h;
  wells = W;
  wi.clear();
  Bp_h.clear();
  Cp_h.clear();
  D_h.clear();
  if (nw > 0) {
    const int np = S.nperf;
    std::vector<double> mob_perf(np);
    // gather the perforated cells' mobilities (tiny)
    for (int j = 0; j < np; ++j)
      PMX_HIP(hipMemcpyAsync(&mob_perf[j], mob + W.cells[j], sizeof(double), hipMemcpyDeviceToHost, stream));
    PMX_HIP(hipStreamSynchronize(stream));
    wi.resize(np);
    std::vector<double> radd(np), Bp(np, 0.0), Cp(np, 0.0), D(nw, 0.0), rw(nw, 0.0);
    for (int w = 0; w < nw; ++w) {
      double wmax = 0.0, wsum = 0.0, wdp = 0.0;
      for (int j = W.perf_ptr[w]; j < W.perf_ptr[w + 1]; ++j) {
        wi[j] = W.WI[j] * mob_perf[j];
        wmax = std::max(wmax, wi[j]);
        wsum += wi[j];
        wdp += wi[j] * W.dp[j];
      }
      for (int j = W.perf_ptr[w]; j < W.perf_ptr[w + 1]; ++j) {
        if (W.type[w] == 0) {
          radd[j] = wi[j] * (W.val[w] + W.dp[j]);
        } else {
          radd[j] = wi[j] * W.dp[j];
          Bp[j] = -wi[j];
          Cp[j] = -wi[j];
        }
      }
      if (W.type[w] == 0) {
        D[w] = wmax;
        rw[w] = wmax * W.val[w];
      } else {
        D[w] = wsum;
        rw[w] = W.val[w] - wdp;
      }
    }
    wi_d.upload(wi, stream);
    radd_d.upload(radd, stream);
    PMX_HIP(hipMemcpyAsync(S.Bp.p, Bp.data(), sizeof(double) * np, hipMemcpyHostToDevice, stream));
    PMX_HIP(hipMemcpyAsync(S.Cp.p, Cp.data(), sizeof(double) * np, hipMemcpyHostToDevice, stream));
    PMX_HIP(hipMemcpyAsync(S.Dp.p, D.data(), sizeof(double) * nw, hipMemcpyHostToDevice, stream));
    Bp_h = Bp;
    Cp_h = Cp;
    D_h = D;
    PMX_HIP(hipMemcpyAsync(r.p + nc, rw.data(), sizeof(double) * nw, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_incomp_perfs, dim3(grid_for(S.npc)), dim3(256), 0, stream, S.npc, S.pc_cell.p, S.pc_ptr.p,
                       S.pc_perf.p, S.diag.p, wi_d.p, radd_d.p, S.val.p, r.p);
    PMX_HIP(hipGetLastError());
    S.has_mix = false;
  }
  // null-space pin (incomp_tpfa.py:355-363): without Dirichlet faces or bhp
  // wells the first diagonal entry is doubled (or the largest one if the
  // first is not positive)
  bool anchored = std::any_of(dF_h.begin(), dF_h.end(), [](unsigned char v) { return v != 0; });
  for (int w = 0; w < nw; ++w) anchored |= W.type[w] == 0;
  if (!anchored && nc > 0) {
    std::vector<double> dg(nc);
    std::vector<double> vals(S.nnzb);
    PMX_HIP(hipMemcpyAsync(vals.data(), S.val.p, sizeof(double) * S.nnzb, hipMemcpyDeviceToHost, stream));
    PMX_HIP(hipStreamSynchronize(stream));
    for (int c = 0; c < nc; ++c) dg[c] = vals[host_pat.diag[c]];
    int k = host_pat.diag[0];
    if (!(dg[0] > 0)) k = host_pat.diag[std::max_element(dg.begin(), dg.end()) - dg.begin()];
    hipLaunchKernelGGL(k_scale_entry, dim3(1), dim3(64), 0, stream, S.val.p, k, 2.0);
    PMX_HIP(hipGetLastError());
  }
  PMX_HIP(hipStreamSynchronize(stream));
  assembled = true;
}

int IncompDev::solve_assembled(const LinearSolverOptions& o, double* x, SolveInfo& info) {
  PMX_CHECK(assembled, "assemble before solving");
  // AMG on the pressure matrix inside FGMRES
  if (!cpr_has_hierarchy(cpr)) {
    std::vector<double> vals(S.nnzb);
    PMX_HIP(hipMemcpyAsync(vals.data(), S.val.p, sizeof(double) * S.nnzb, hipMemcpyDeviceToHost, stream));
    PMX_HIP(hipStreamSynchronize(stream));
    std::vector<double> wts(S.nnzb, 0.0);
    for (int c = 0; c < nc; ++c)
      for (int k = host_pat.rowptr[c]; k < host_pat.rowptr[c + 1]; ++k)
        if (host_pat.col[k] != c) wts[k] = std::fabs(vals[k]);
    cpr_build_hierarchy(cpr, S, wts, o, stream);
  }
  cpr_setup(cpr, S, w1.p, o, stream);
  return cpr_solve(cpr, S, r.p, x, o, info, stream);
}

// Pressures, face pressures, fluxes and well quantities from a solution x
// (device: cell pressures, then the pressure of each well) of the system
// last assembled (incomp_tpfa.py:365-389).
void IncompDev::reconstruct(const double* x, double* p, double* flux, double* fpress, double* well_p,
                            double* well_flux) {
  PMX_CHECK(assembled, "assemble before reconstructing");
  const IncompWells& W = wells;
  const int nw = (int)W.type.size();
  PMX_HIP(hipMemcpyAsync(p, x, sizeof(double) * nc, hipMemcpyDeviceToDevice, stream));
  hipLaunchKernelGGL(k_incomp_reconstruct, dim3(grid_for(nf)), dim3(256), 0, stream, nf, face_hf.p, face_nb.p,
                     hf_cell.p, thf.p, ghf.p, ft.p, fg.p, fg_all.p, dF.p, dval.p, hh.p, x, fpress, flux);
  PMX_HIP(hipGetLastError());
  if (nw > 0) {
    std::vector<double> x_h(S.n());
    PMX_HIP(hipMemcpyAsync(x_h.data(), x, sizeof(double) * S.n(), hipMemcpyDeviceToHost, stream));
    PMX_HIP(hipStreamSynchronize(stream));
    for (int w = 0; w < nw; ++w) {
      well_p[w] = x_h[nc + w];
      for (int j = W.perf_ptr[w]; j < W.perf_ptr[w + 1]; ++j)
        well_flux[j] = wi_h[j] * (x_h[nc + w] + W.dp[j] - x_h[W.cells[j]]);
    }
  }
  PMX_HIP(hipStreamSynchronize(stream));
}

int IncompDev::solve(const double* T, const double* mob, const double* omega, const double* g, const IncompBC& bc,
                     const IncompWells& W, const LinearSolverOptions& o, double* p, double* flux, double* fpress,
                     double* well_p, double* well_flux, SolveInfo& info) {
  assemble(T, mob, omega, g, bc, W);
  const int st = solve_assembled(o, x_work(), info);
  reconstruct(xw.p, p, flux, fpress, well_p, well_flux);
  return st;
}

// The assembled system on the host: the cell matrix in CSR (rowptr[nc+1],
// col, val), per perforation the cell-row / well-column and well-row /
// cell-column entries (Bp, Cp), the well diagonal D[nw], and rhs[nc+nw].
void IncompDev::export_system(int* rowptr, int* col, double* val, double* Bp, double* Cp, double* D, double* rhs) {
  PMX_CHECK(assembled, "assemble before exporting the system");
  if (rowptr) std::copy(host_pat.rowptr.begin(), host_pat.rowptr.end(), rowptr);
  if (col) std::copy(host_pat.col.begin(), host_pat.col.end(), col);
  if (val) PMX_HIP(hipMemcpyAsync(val, S.val.p, sizeof(double) * S.nnzb, hipMemcpyDeviceToHost, stream));
  if (rhs) PMX_HIP(hipMemcpyAsync(rhs, r.p, sizeof(double) * S.n(), hipMemcpyDeviceToHost, stream));
  PMX_HIP(hipStreamSynchronize(stream));
  if (Bp) std::copy(Bp_h.begin(), Bp_h.end(), Bp);
  if (Cp) std::copy(Cp_h.begin(), Cp_h.end(), Cp);
  if (D) std::copy(D_h.begin(), D_h.end(), D);
}

double* IncompDev::x_work() {
  xw.alloc(S.n());
  return xw.p;
}

}  // namespace pmx
