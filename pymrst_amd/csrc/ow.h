// Device oil-water model (see ow.hip).
#pragma once
#include "linsys.h"

namespace pmx {

// init_simple_adi_fluid for phases 'WO': b_a = exp(c_a (p - p_ref)) (1 when
// c_a == 0), kr_w = sW^n_w, kr_o = (1-sW)^n_o, constant viscosities; with
// has_cr the pore volume is multiplied by 1 + c_r (p - p_ref).
struct OWFluid {
  double mu_w, mu_o, rho_w, rho_o, n_w, n_o, c_w, c_o, p_ref, c_r;
  int has_cr;
};

class OWModelDev {
 public:
  explicit OWModelDev(hipStream_t s);
  ~OWModelDev();
  void set_topology(int nc, int nconn, const int* N, const double* T, const double* gdz, const double* pv);
  void set_wells(int nw, const int* perf_ptr, const int* perf_cells, const double* WI, const int* types,
                 const double* vals, const double* compi);
  void assemble(const double* p, const double* s, const double* qws, const double* qos, const double* bhp,
                const double* p0, const double* s0, double dt, const double* cdp_dev);
  // controls as linear forms of the well unknowns (coef[3 w..] on qWs, qOs,
  // bhp; val the target), open/shut status and the rate-row scales
  void set_controls(const double* coef, const double* vals, const int* status, const double* rscale);
  void convergence(const double* p, double* out7, double* wells_out);
  // the same with the current pore volume and its sum as an eighth value
  void convergence_current_pv(const double* p, double* out8, double* wells_out);
  // the same reductions with the partials in a static per-GPU array: no
  // allocation on any call (device: the GPU this model runs on, current)
  void convergence_noalloc(int device, const double* p, double* out7, double* wells_out);
  void reduce_convergence(const double* p, double* out, int nout, double* wells_out, double* partial, int nblk,
                          int cur_pv);
  int solve(const LinearSolverOptions& o, double* dx, SolveInfo& info, double pmean);
  void update(const double* dx, double sign, double ds_max, double dp_max_rel, double* p, double* s, double* qws,
              double* qos, double* bhp);
  void mask_residual(const double* mask);   // r *= mask (device, per unknown)
  // members of a union model: consecutive cell ranges [cell_ptr[m],
  // cell_ptr[m+1]); member_convergence reduces each member's cells (out:
  // nmem rows of max|r_w/pv|, max|r_o/pv|, sum r_w, sum r_o, sum 1/b_w(p),
  // sum 1/b_o(p), sum p), in a fixed order
  void set_members(int nmem, const int* cell_ptr);
  void member_convergence(const double* p, double* out7, double* wells_out);
  const double* dt_cell = nullptr;          // per-cell step length (device, caller-owned), or the scalar dt

  // adjoint (adjoint.hip): transposed solve of the last assembled Jacobian,
  // the previous-state coupling and the log-permeability derivative
  void set_trans_fractions(const double* frac);
  void transpose_system();
  int adjoint_solve(const LinearSolverOptions& o, double pmean, const double* rhs, double* lam, SolveInfo& info);
  void prev_state_transpose(const double* p0, const double* s0, double dt, const double* lam, double* out);
  void dlogk(const double* p, const double* s, const double* qws, const double* qos, const double* bhp,
             const double* cdp_dev, const double* lam, double* grad);

  hipStream_t stream;
  OWFluid fluid{};
  int nc = 0, nconn = 0, nw = 0;
  BlockSystem S;
  DArr<int> bconn;
  DArr<double> bT, bG, pv;
  DArr<double> WI, cdp, wval, compi, rq, dq, wctrl, qvol;
  DArr<int> wstat, mixflag;
  double rscale[2] = {1.0, 1.0};
  DArr<double> r, rhs, w_cpr, conv_partial;
  int nmem = 0, nchunk = 0;
  DArr<int> mem_chunk, mem_chunk_ptr;      // per chunk (i0, i1); chunks of member m: [ptr[m], ptr[m+1])
  DArr<double> mem_partial, mem_out;
  std::vector<double> hier_weights;
  std::vector<int> host_bconn, line_ptr, line_cells;
  CprSolver* cpr = nullptr;
  // adjoint
  BlockSystem ST;
  DArr<int> tmap;                      // block k = (i,j) -> block (j,i)
  DArr<double> bfrac, w_adj, gperf;    // per block: T_f / T_h of the row's half-face
  CprSolver* cprT = nullptr;
};

}  // namespace pmx
