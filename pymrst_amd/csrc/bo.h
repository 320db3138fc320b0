// Device three-phase live-oil black-oil model (see bo.hip).
#pragma once
#include "linsys.h"

namespace pmx {

// Deck black-oil fluid (init_deck_blackoil_fluid) as flat tables.  All
// offsets index one table buffer; the PVTO branch rows of key k are rows
// [bptr[k], bptr[k+1]) of the br_* columns.  Every table is ascending in its
// first column.  The functions reading it are __host__ __device__ so the host
// evaluates the same interpolation for the CPR weights.
struct BOFluidView {
  const double* t = nullptr;
  const int* bptr = nullptr;
  int n_rs = 0, rs_p = 0, rs_k = 0;          // rs_sat(p): (0, 0) then (p_sat, rs) per PVTO key
  int n_g = 0, g_p = 0, g_b = 0, g_mu = 0;   // PVDG with a leading row at p = 0: b_g = 1/Bg, mu_g
  int n_w = 0, w_s = 0, w_kr = 0;            // SWOF extended: kr_w(sw)
  int n_ow = 0, ow_s = 0, ow_kr = 0;         // 1 - sw reversed: kr_ow(so)
  int n_gg = 0, gg_s = 0, gg_kr = 0;         // SGOF extended: kr_g(sg)
  int n_og = 0, og_s = 0, og_kr = 0;         // 1 - (sg + swcon) reversed: kr_og(so)
  int nkey = 0, key = 0, sat_p = 0, sat_b = 0, sat_mu = 0;   // PVTO saturated curve per rs key
  int br_dp = 0, br_b = 0, br_mu = 0;        // branch rows: p - p_sat, 1/Bo, mu_o
  double rhoW = 0, rhoO = 0, rhoG = 0;
  double pw_ref = 0, bw_ref = 1, cw = 0, muw = 1;
  double pr_ref = 0, cr = 0;
  double swcon = 0;
};

// Host-side owner of the tables; view() points at the device copy,
// host_view() at the host copy.
struct BOFluidTables {
  std::vector<double> tab;
  std::vector<int> bptr;
  BOFluidView layout;   // offsets and scalars (pointers unset)
  DArr<double> dtab;
  DArr<int> dbptr;
  void build(const double* scalars, int n_pvdg, const double* pvdg, int n_key, const double* keys,
             const int* branch_ptr, const double* branch, int n_swof, const double* swof, int n_sgof,
             const double* sgof);
  void upload(hipStream_t s);
  BOFluidView view() const {
    BOFluidView v = layout;
    v.t = dtab.p;
    v.bptr = dbptr.p;
    return v;
  }
  BOFluidView host_view() const {
    BOFluidView v = layout;
    v.t = tab.data();
    v.bptr = bptr.data();
    return v;
  }
};

class BOModelDev {
 public:
  explicit BOModelDev(hipStream_t s);
  ~BOModelDev();
  void set_topology(int nc, int nconn, const int* N, const double* T, const double* gdz, const double* pv);
  void set_wells(int nw, const int* perf_ptr, const int* perf_cells, const double* WI, const int* types,
                 const double* vals, const double* compi);
  // state arrays: p[nc], s[3 nc] (sw, so, sg per cell), rs[nc], status[nc]
  // (1 undersaturated, 3 saturated), q[3 nw] (qWs, qOs, qGs per well), bhp[nw]
  void begin_step(const double* p0, const double* s0, const double* rs0, const int* status0);
  void assemble(const double* p, const double* s, const double* rs, const int* status, const double* q,
                const double* bhp, double dt);
  // out12: max|r_a/pv| (3), sum r_a (3), sum 1/b_a (3), sum p, sum rs, sum pv
  void convergence(const double* p, const double* rs, const int* status, double* out12, double* wells_out);
  int solve(const LinearSolverOptions& o, double* dx, SolveInfo& info, double pmean, double rsmean);
  void update(const double* dx, double ds_max, double* p, double* s, double* rs, int* status, double* q,
              double* bhp);
  // the facility's well pieces (GenericBlackOilModelAD's wells): per-well
  // control rows c_w qWs + c_o qOs + c_g qGs + c_b bhp - val (coef[4 nw]),
  // open/shut status, the rate rows' per-phase scale; per-perforation
  // connection drops; the update rule (0 hand-rolled, 1 generic)
  void set_controls(const double* coef, const double* val, const int* status, const double* rscale);
  void set_connection_drops(const double* cdp_host);
  void perf_volume_rates(double* out_host);
  void set_update_rule(int rule, double dp_max_rel, double drs_max);
  // per cell (rho_w, rho_o, rho_g, lambda_w, lambda_o, lambda_g) into out[6 nc] (device)
  void cell_densities(const double* p, const double* s, const double* rs, const int* status, double* out);

  hipStream_t stream;
  BOFluidTables fluid;
  int nc = 0, nconn = 0, nw = 0;
  BlockSystem S;
  DArr<int> bconn;
  DArr<double> bT, bG, pv, props, acc0;
  DArr<double> WI, wval, compi, rq, dq, cdp, wctrl, qvol;
  DArr<int> wtype, mixflag, wstat;
  double rsc[3] = {1.0, 1.0, 1.0};
  int upd_rule = 0;
  double dp_max_rel = 0.0, drs_max = 0.0;
  DArr<double> r, rhs, w_cpr, conv_partial;
  std::vector<double> hier_weights;
  std::vector<int> line_ptr, line_cells;
  CprSolver* cpr = nullptr;
};

}  // namespace pmx
