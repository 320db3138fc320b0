// C ABI entry points (include/pmx.h).  Exceptions never cross the boundary:
// each entry point converts them into a negative status and a message.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <string>

#include "../../include/pmx.h"
#include "incomp.h"
#include "ow.h"
#include "bo.h"
#include "cpgrid.h"

using namespace pmx;

struct pmx_incomp {
  int device = 0;
  std::unique_ptr<IncompDev> m;
};

struct pmx_ow {
  int device = 0;
  std::unique_ptr<OWModelDev> m;
  pmx_solver_options last_opts{};
};

struct pmx_cp {
  CpGrid g;
};

struct pmx_bo {
  int device = 0;
  std::unique_ptr<BOModelDev> m;
  pmx_solver_options last_opts{};
};

// A block system given by its caller: cell blocks in BCSR (any pattern that
// holds every diagonal block), per-perforation well blocks.
struct pmx_sys {
  int device = 0;
  hipStream_t stream = nullptr;
  BlockSystem S;
  CprSolver* cpr = nullptr;
  DArr<double> w;                 // pressure weights (B)
  std::vector<int> h_rowptr, h_col;
  bool have_values = false;
  ~pmx_sys() {
    if (cpr) cpr_destroy(cpr);
  }
};

static thread_local std::string g_err;

#define PMX_TRY(body)                       \
  try {                                     \
    body;                                   \
    return 0;                               \
  } catch (const std::exception& e) {       \
    g_err = e.what();                       \
    return -1;                              \
  } catch (...) {                           \
    g_err = "unknown error";                \
    return -1;                              \
  }

static LinearSolverOptions to_opts(const pmx_solver_options* o) {
  LinearSolverOptions r;
  if (!o) return r;
  r.rtol = o->rtol;
  r.restart = o->restart;
  r.maxiter = o->maxiter;
  r.precision = o->precision;
  r.use_pressure_stage = o->use_pressure_stage;
  r.pre_smooth = o->pre_smooth;
  r.post_smooth = o->post_smooth;
  r.jacobi_omega = o->jacobi_omega;
  r.coarse_alpha = o->coarse_alpha;
  r.theta = o->theta;
  r.coarse_size = o->coarse_size;
  r.pairwise_passes = o->pairwise_passes;
  r.reorthogonalize = o->reorthogonalize;
  r.verbose = o->verbose;
  r.local_stage = o->local_stage;
  r.amg_cycles = o->pressure_cycles;
  r.local_sweeps = o->local_sweeps;
  r.use_graphs = o->use_graphs;
  r.coarse_block_rows = o->coarse_block_rows;
  r.line_segments = o->line_segments;
  r.autotune = o->autotune;
  r.level_shape = o->level_shape;
  r.dot_shape = o->dot_shape;
  return r;
}

static void bind(pmx_ow* h) { PMX_HIP(hipSetDevice(h->device)); }
static void bind(pmx_bo* h) { PMX_HIP(hipSetDevice(h->device)); }

static void fill_info(pmx_solve_info* info, const SolveInfo& si) {
  if (!info) return;
  info->iterations = si.iterations;
  info->restarts = si.restarts;
  info->rel_residual = si.rel_residual;
  info->setup_ms = si.setup_ms;
  info->solve_ms = si.solve_ms;
  info->levels = si.levels;
  info->coarse_n = si.coarse_n;
}

extern "C" {

const char* pmx_last_error(void) { return g_err.c_str(); }

void pmx_solver_defaults(pmx_solver_options* o) {
  LinearSolverOptions d;
  o->rtol = d.rtol;
  o->restart = d.restart;
  o->maxiter = d.maxiter;
  o->precision = d.precision;
  o->use_pressure_stage = d.use_pressure_stage;
  o->pre_smooth = d.pre_smooth;
  o->post_smooth = d.post_smooth;
  o->jacobi_omega = d.jacobi_omega;
  o->coarse_alpha = d.coarse_alpha;
  o->theta = d.theta;
  o->coarse_size = d.coarse_size;
  o->pairwise_passes = d.pairwise_passes;
  o->reorthogonalize = d.reorthogonalize;
  o->verbose = d.verbose;
  o->local_stage = d.local_stage;
  o->pressure_cycles = d.amg_cycles;
  o->local_sweeps = d.local_sweeps;
  o->use_graphs = d.use_graphs;
  o->coarse_block_rows = d.coarse_block_rows;
  o->line_segments = d.line_segments;
  o->autotune = d.autotune;
  o->level_shape = d.level_shape;
  o->dot_shape = d.dot_shape;
}

int pmx_ow_create(int32_t device, void* hip_stream, const double* fluid, pmx_ow** out) {
  PMX_TRY({
    PMX_CHECK(out != nullptr && fluid != nullptr, "null argument");
    auto h = std::make_unique<pmx_ow>();
    h->device = device;
    bind(h.get());
    h->m = std::make_unique<OWModelDev>((hipStream_t)hip_stream);
    OWFluid& f = h->m->fluid;
    f.mu_w = fluid[0];
    f.mu_o = fluid[1];
    f.rho_w = fluid[2];
    f.rho_o = fluid[3];
    f.n_w = fluid[4];
    f.n_o = fluid[5];
    f.c_w = fluid[6];
    f.c_o = fluid[7];
    f.p_ref = fluid[8];
    f.c_r = fluid[9] >= 0 ? fluid[9] : 0.0;
    f.has_cr = fluid[9] >= 0 ? 1 : 0;
    pmx_solver_defaults(&h->last_opts);
    *out = h.release();
  })
}

int pmx_ow_destroy(pmx_ow* h) {
  PMX_TRY({
    if (h) {
      bind(h);
      delete h;
    }
  })
}

int pmx_ow_set_topology(pmx_ow* h, int32_t nc, int32_t nconn, const int32_t* N, const double* T, const double* gdz,
                        const double* pv) {
  PMX_TRY({
    PMX_CHECK(h && N && T && gdz && pv, "null argument");
    bind(h);
    h->m->set_topology(nc, nconn, N, T, gdz, pv);
  })
}

int pmx_ow_set_wells(pmx_ow* h, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells, const double* WI,
                     const int32_t* type, const double* val, const double* compi) {
  PMX_TRY({
    PMX_CHECK(h && (nw == 0 || (perf_ptr && perf_cells && WI && type && val && compi)), "null argument");
    bind(h);
    for (int w = 0; w < nw; ++w) PMX_CHECK(type[w] == 0 || type[w] == 1, "unsupported well control");
    h->m->set_wells(nw, perf_ptr, perf_cells, WI, type, val, compi);
  })
}

int pmx_ow_assemble(pmx_ow* h, const double* p, const double* sw, const double* qws, const double* qos,
                    const double* bhp, const double* p0, const double* sw0, double dt, const double* cdp) {
  PMX_TRY({
    PMX_CHECK(h && p && sw && p0 && sw0, "null argument");
    PMX_CHECK(dt > 0, "dt must be positive");
    bind(h);
    h->m->assemble(p, sw, qws, qos, bhp, p0, sw0, dt, cdp);
  })
}

int pmx_ow_residual(pmx_ow* h, double* out) {
  PMX_TRY({
    PMX_CHECK(h && out, "null argument");
    bind(h);
    PMX_HIP(hipMemcpyAsync(out, h->m->r.p, sizeof(double) * h->m->S.n(), hipMemcpyDeviceToDevice, h->m->stream));
  })
}

int pmx_ow_convergence(pmx_ow* h, const double* p, double* out7, double* wells_out) {
  PMX_TRY({
    PMX_CHECK(h && p && out7, "null argument");
    bind(h);
    h->m->convergence(p, out7, wells_out);
  })
}

int pmx_ow_set_controls(pmx_ow* h, const double* coef, const double* val, const int32_t* status,
                        const double* rscale) {
  PMX_TRY({
    PMX_CHECK(h && rscale, "null argument");
    bind(h);
    auto& m = *h->m;
    PMX_CHECK(m.nw == 0 || (coef && val && status), "null well control argument");
    for (int w = 0; w < m.nw; ++w) PMX_CHECK(status[w] == 0 || status[w] == 1, "well status is 0 (shut) or 1");
    PMX_CHECK(std::isfinite(rscale[0]) && std::isfinite(rscale[1]) && rscale[0] > 0 && rscale[1] > 0,
              "rate-row scales must be positive");
    m.set_controls(coef, val, status, rscale);
  })
}

int pmx_ow_convergence_current_pv(pmx_ow* h, const double* p, double* out8, double* wells_out) {
  PMX_TRY({
    PMX_CHECK(h && p && out8, "null argument");
    bind(h);
    h->m->convergence_current_pv(p, out8, wells_out);
  })
}

int pmx_ow_perf_volume_rates(pmx_ow* h, double* out) {
  PMX_TRY({
    PMX_CHECK(h && out, "null argument");
    bind(h);
    auto& m = *h->m;
    if (m.S.nperf > 0) {
      PMX_HIP(hipMemcpyAsync(out, m.qvol.p, sizeof(double) * 2 * m.S.nperf, hipMemcpyDeviceToHost, m.stream));
      PMX_HIP(hipStreamSynchronize(m.stream));
    }
  })
}

int pmx_ow_convergence_noalloc(pmx_ow* h, const double* p, double* out7, double* wells_out) {
  PMX_TRY({
    PMX_CHECK(h && p && out7, "null argument");
    bind(h);
    h->m->convergence_noalloc(h->device, p, out7, wells_out);
  })
}

int pmx_ow_solve(pmx_ow* h, const pmx_solver_options* o, double pmean, double* dx, pmx_solve_info* info) {
  try {
    PMX_CHECK(h && dx, "null argument");
    bind(h);
    SolveInfo si;
    LinearSolverOptions lo = to_opts(o);
    if (o) h->last_opts = *o;
    const int st = h->m->solve(lo, dx, si, pmean);
    if (info) {
      info->iterations = si.iterations;
      info->restarts = si.restarts;
      info->rel_residual = si.rel_residual;
      info->setup_ms = si.setup_ms;
      info->solve_ms = si.solve_ms;
      info->levels = si.levels;
      info->coarse_n = si.coarse_n;
    }
    if (st < 0) {
      g_err = "linear solve produced a non-finite residual";
      return -2;
    }
    return st;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int pmx_ow_update(pmx_ow* h, const double* dx, double sign, double ds_max, double dp_max_rel, double* p, double* sw,
                  double* qws, double* qos, double* bhp) {
  PMX_TRY({
    PMX_CHECK(h && dx && p && sw, "null argument");
    bind(h);
    h->m->update(dx, sign, ds_max, dp_max_rel, p, sw, qws, qos, bhp);
  })
}

int pmx_ow_matvec(pmx_ow* h, const double* x, double* y) {
  PMX_TRY({
    PMX_CHECK(h && x && y, "null argument");
    bind(h);
    DArr<double> hs;
    hs.alloc((size_t)std::max(h->m->nw, 1) * kMaxMix);
    system_matvec(h->m->S, x, y, hs.p, h->m->stream);
    PMX_HIP(hipStreamSynchronize(h->m->stream));
  })
}

int pmx_ow_nnzb(pmx_ow* h, int32_t* nnzb) {
  PMX_TRY({
    PMX_CHECK(h && nnzb, "null argument");
    *nnzb = h->m->S.nnzb;
  })
}

int pmx_ow_export_pattern(pmx_ow* h, int32_t* rowptr, int32_t* col) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    bind(h);
    auto& S = h->m->S;
    if (rowptr)
      PMX_HIP(hipMemcpyAsync(rowptr, S.rowptr.p, sizeof(int) * (S.nc + 1), hipMemcpyDeviceToHost, h->m->stream));
    if (col) PMX_HIP(hipMemcpyAsync(col, S.col.p, sizeof(int) * S.nnzb, hipMemcpyDeviceToHost, h->m->stream));
    PMX_HIP(hipStreamSynchronize(h->m->stream));
  })
}

int pmx_ow_export_blocks(pmx_ow* h, double* vals) {
  PMX_TRY({
    PMX_CHECK(h && vals, "null argument");
    bind(h);
    auto& S = h->m->S;
    PMX_HIP(hipMemcpyAsync(vals, S.val.p, sizeof(double) * 4 * S.nnzb, hipMemcpyDeviceToDevice, h->m->stream));
  })
}

int pmx_ow_connection_rates(pmx_ow* h, double* out) {
  PMX_TRY({
    PMX_CHECK(h && out, "null argument");
    bind(h);
    auto& m = *h->m;
    if (m.S.nperf > 0) {
      PMX_HIP(hipMemcpyAsync(out, m.rq.p, sizeof(double) * 2 * m.S.nperf, hipMemcpyDeviceToHost, m.stream));
      PMX_HIP(hipStreamSynchronize(m.stream));
    }
  })
}

int pmx_ow_set_lines(pmx_ow* h, int32_t nlines, const int32_t* line_ptr, const int32_t* cells) {
  PMX_TRY({
    PMX_CHECK(h && line_ptr && cells && nlines >= 0, "null argument");
    cpr_set_lines(h->m->cpr, nlines, line_ptr, cells);
    h->m->line_ptr.assign(line_ptr, line_ptr + nlines + 1);
    h->m->line_cells.assign(cells, cells + line_ptr[nlines]);
  })
}

int pmx_ow_profile(pmx_ow* h, int32_t enable) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    cpr_profile(h->m->cpr, enable != 0);
  })
}

int pmx_ow_profile_read(pmx_ow* h, double* out4) {
  PMX_TRY({
    PMX_CHECK(h && out4, "null argument");
    bind(h);
    cpr_profile_read(h->m->cpr, out4);
  })
}

int pmx_ow_launch_shapes(pmx_ow* h, int32_t* out, int32_t cap, int32_t* count) {
  PMX_TRY({
    PMX_CHECK(h && out && count && cap >= 2, "bad argument");
    *count = cpr_launch_shapes(h->m->cpr, h->m->S, out, cap);
  })
}

int pmx_autotune_reset(void) {
  PMX_TRY({ tune_table_clear(); })
}

int pmx_autotune_save(const char* path, int32_t* count) {
  PMX_TRY({
    PMX_CHECK(path != nullptr, "null path");
    const int n = tune_table_save(path);
    if (count) *count = n;
  })
}

int pmx_autotune_load(const char* path, int32_t* count) {
  PMX_TRY({
    PMX_CHECK(path != nullptr, "null path");
    const int n = tune_table_load(path);
    if (count) *count = n;
  })
}

int pmx_autotune_entries(int32_t* count) {
  PMX_TRY({
    PMX_CHECK(count != nullptr, "null argument");
    *count = tune_table_size();
  })
}

int pmx_ow_profile_line_gs(pmx_ow* h, double* out3) {
  PMX_TRY({
    PMX_CHECK(h && out3, "null argument");
    bind(h);
    cpr_profile_line_read(h->m->cpr, out3);
  })
}

int pmx_incomp_create(int32_t device, void* hip_stream, pmx_incomp** out) {
  PMX_TRY({
    PMX_CHECK(out != nullptr, "null argument");
    auto h = std::make_unique<pmx_incomp>();
    h->device = device;
    PMX_HIP(hipSetDevice(device));
    h->m = std::make_unique<IncompDev>((hipStream_t)hip_stream);
    *out = h.release();
  })
}

int pmx_incomp_destroy(pmx_incomp* h) {
  PMX_TRY({
    if (h) {
      PMX_HIP(hipSetDevice(h->device));
      delete h;
    }
  })
}

int pmx_incomp_set_grid(pmx_incomp* h, int32_t nc, int32_t nf, const int32_t* face_pos, const int32_t* hf_face,
                        const int32_t* face_nb, const double* cell_centroids, const double* face_centroids) {
  PMX_TRY({
    PMX_CHECK(h && face_pos && hf_face && face_nb && cell_centroids && face_centroids, "null argument");
    PMX_HIP(hipSetDevice(h->device));
    h->m->set_grid(nc, nf, face_pos, hf_face, face_nb, cell_centroids, face_centroids);
  })
}

int pmx_compute_trans(pmx_incomp* h, const double* face_normals, int32_t nk, const double* perm, double* T) {
  PMX_TRY({
    PMX_CHECK(h && face_normals && perm && T, "null argument");
    PMX_CHECK(nk == 1 || nk == 3 || nk == 6, "permeability must have 1, 3 or 6 columns");
    PMX_HIP(hipSetDevice(h->device));
    h->m->compute_trans(face_normals, nk, perm, T);
  })
}

static IncompBC incomp_bc(int32_t nbc, const int32_t* bc_face, const int32_t* bc_type, const double* bc_value,
                           int32_t nsrc, const int32_t* src_cell, const double* src_rate) {
  IncompBC bc;
  if (nbc > 0) {
    PMX_CHECK(bc_face && bc_type && bc_value, "null boundary-condition argument");
    bc.face.assign(bc_face, bc_face + nbc);
    bc.type.assign(bc_type, bc_type + nbc);
    bc.value.assign(bc_value, bc_value + nbc);
  }
  if (nsrc > 0) {
    PMX_CHECK(src_cell && src_rate, "null source argument");
    bc.src_cell.assign(src_cell, src_cell + nsrc);
    bc.src_rate.assign(src_rate, src_rate + nsrc);
  }
  return bc;
}

static IncompWells incomp_wells(const pmx_incomp* h, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells,
                                const double* WI, const int32_t* wtype, const double* wval, const double* wdp) {
  IncompWells W;
  if (nw > 0) {
    PMX_CHECK(perf_ptr && perf_cells && WI && wtype && wval && wdp, "null well argument");
    const int np = perf_ptr[nw];
    W.perf_ptr.assign(perf_ptr, perf_ptr + nw + 1);
    W.cells.assign(perf_cells, perf_cells + np);
    W.WI.assign(WI, WI + np);
    W.dp.assign(wdp, wdp + np);
    W.type.assign(wtype, wtype + nw);
    W.val.assign(wval, wval + nw);
    for (int c : W.cells) PMX_CHECK(c >= 0 && c < h->m->nc, "well connection cell out of range");
    for (int t : W.type) PMX_CHECK(t == 0 || t == 1, "unsupported well control");
  }
  return W;
}

int pmx_incomp_solve(pmx_incomp* h, const double* T, const double* mob, const double* omega, const double* gravity,
                     int32_t nbc, const int32_t* bc_face, const int32_t* bc_type, const double* bc_value, int32_t nsrc,
                     const int32_t* src_cell, const double* src_rate, int32_t nw, const int32_t* perf_ptr,
                     const int32_t* perf_cells, const double* WI, const int32_t* wtype, const double* wval,
                     const double* wdp, const pmx_solver_options* o, double* p, double* flux, double* face_pressure,
                     double* well_p, double* well_flux, pmx_solve_info* info) {
  try {
    PMX_CHECK(h && T && mob && omega && p && flux && face_pressure, "null argument");
    PMX_CHECK(nw == 0 || (well_p && well_flux), "null well output");
    PMX_HIP(hipSetDevice(h->device));
    const IncompBC bc = incomp_bc(nbc, bc_face, bc_type, bc_value, nsrc, src_cell, src_rate);
    const IncompWells W = incomp_wells(h, nw, perf_ptr, perf_cells, WI, wtype, wval, wdp);
    SolveInfo si;
    const int st = h->m->solve(T, mob, omega, gravity, bc, W, to_opts(o), p, flux, face_pressure, well_p,
                               well_flux, si);
    fill_info(info, si);
    if (st < 0) {
      g_err = "pressure solve produced a non-finite residual";
      return -2;
    }
    return st;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int pmx_incomp_assemble(pmx_incomp* h, const double* T, const double* mob, const double* omega,
                        const double* gravity, int32_t nbc, const int32_t* bc_face, const int32_t* bc_type,
                        const double* bc_value, int32_t nsrc, const int32_t* src_cell, const double* src_rate,
                        int32_t nbcp, const int32_t* bcp_face, const double* bcp_value, int32_t nw,
                        const int32_t* perf_ptr, const int32_t* perf_cells, const double* WI, const int32_t* wtype,
                        const double* wval, const double* wdp) {
  PMX_TRY({
    PMX_CHECK(h && T && mob && omega, "null argument");
    PMX_HIP(hipSetDevice(h->device));
    IncompBC bc = incomp_bc(nbc, bc_face, bc_type, bc_value, nsrc, src_cell, src_rate);
    if (nbcp > 0) {
      PMX_CHECK(bcp_face && bcp_value, "null periodic boundary-condition argument");
      bc.jump_face.assign(bcp_face, bcp_face + nbcp);
      bc.jump_value.assign(bcp_value, bcp_value + nbcp);
    }
    h->m->assemble(T, mob, omega, gravity, bc, incomp_wells(h, nw, perf_ptr, perf_cells, WI, wtype, wval, wdp));
  })
}

int pmx_incomp_nnz(pmx_incomp* h, int32_t* nnz) {
  PMX_TRY({
    PMX_CHECK(h && nnz, "null argument");
    *nnz = h->m->S.nnzb;
  })
}

int pmx_incomp_export(pmx_incomp* h, int32_t* rowptr, int32_t* col, double* val, double* Bp, double* Cp, double* D,
                      double* rhs) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    PMX_HIP(hipSetDevice(h->device));
    h->m->export_system(rowptr, col, val, Bp, Cp, D, rhs);
  })
}

int pmx_incomp_solve_assembled(pmx_incomp* h, const pmx_solver_options* o, double* x, pmx_solve_info* info) {
  try {
    PMX_CHECK(h && x, "null argument");
    PMX_HIP(hipSetDevice(h->device));
    SolveInfo si;
    const int st = h->m->solve_assembled(to_opts(o), x, si);
    fill_info(info, si);
    if (st < 0) {
      g_err = "pressure solve produced a non-finite residual";
      return -2;
    }
    return st;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int pmx_incomp_reconstruct(pmx_incomp* h, const double* x, double* p, double* flux, double* face_pressure,
                           double* well_p, double* well_flux) {
  PMX_TRY({
    PMX_CHECK(h && x && p && flux && face_pressure, "null argument");
    PMX_CHECK(h->m->wells.type.empty() || (well_p && well_flux), "null well output");
    PMX_HIP(hipSetDevice(h->device));
    h->m->reconstruct(x, p, flux, face_pressure, well_p, well_flux);
  })
}

int pmx_ow_precondition(pmx_ow* h, const pmx_solver_options* o, const double* r, double* z) {
  PMX_TRY({
    PMX_CHECK(h && r && z, "null argument");
    bind(h);
    LinearSolverOptions lo = to_opts(o ? o : &h->last_opts);
    cpr_apply(h->m->cpr, h->m->S, r, z, lo, h->m->stream);
  })
}

int pmx_ow_set_trans_fractions(pmx_ow* h, const double* frac) {
  PMX_TRY({
    PMX_CHECK(h && frac, "null argument");
    bind(h);
    h->m->set_trans_fractions(frac);
  })
}

int pmx_ow_adjoint_solve(pmx_ow* h, const pmx_solver_options* o, double pmean, const double* rhs, double* lam,
                         pmx_solve_info* info) {
  try {
    PMX_CHECK(h && rhs && lam, "null argument");
    bind(h);
    LinearSolverOptions lo = to_opts(o ? o : &h->last_opts);
    SolveInfo si;
    const int st = h->m->adjoint_solve(lo, pmean, rhs, lam, si);
    if (info) {
      info->iterations = si.iterations;
      info->restarts = si.restarts;
      info->rel_residual = si.rel_residual;
      info->setup_ms = si.setup_ms;
      info->solve_ms = si.solve_ms;
      info->levels = si.levels;
      info->coarse_n = si.coarse_n;
    }
    if (st < 0) {
      g_err = "adjoint solve produced a non-finite residual";
      return -2;
    }
    return st;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int pmx_ow_prev_state_transpose(pmx_ow* h, const double* p0, const double* s0, double dt, const double* lam,
                                double* out) {
  PMX_TRY({
    PMX_CHECK(h && p0 && s0 && lam && out && dt > 0.0, "bad argument");
    bind(h);
    h->m->prev_state_transpose(p0, s0, dt, lam, out);
  })
}

int pmx_ow_dlogk(pmx_ow* h, const double* p, const double* s, const double* qws, const double* qos,
                 const double* bhp, const double* cdp, const double* lam, double* grad) {
  PMX_TRY({
    PMX_CHECK(h && p && s && lam && grad, "null argument");
    PMX_CHECK(h->m->nw == 0 || (qws && qos && bhp), "well state required");
    bind(h);
    h->m->dlogk(p, s, qws, qos, bhp, cdp, lam, grad);
  })
}

int pmx_ow_set_segments(pmx_ow* h, int32_t nseg, const int32_t* cell_ptr, const int32_t* well_ptr) {
  PMX_TRY({
    PMX_CHECK(h && cell_ptr && well_ptr && nseg >= 1, "bad argument");
    bind(h);
    cpr_set_segments(h->m->cpr, nseg, cell_ptr, well_ptr, 2, 3, h->m->nc, h->m->nw);
  })
}

int pmx_ow_set_cell_dt(pmx_ow* h, const double* dt_cell) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    h->m->dt_cell = dt_cell;
  })
}

int pmx_ow_set_members(pmx_ow* h, int32_t nmem, const int32_t* cell_ptr) {
  PMX_TRY({
    PMX_CHECK(h && cell_ptr && nmem >= 1, "bad argument");
    bind(h);
    h->m->set_members(nmem, cell_ptr);
  })
}

int pmx_ow_member_convergence(pmx_ow* h, const double* p, double* out7, double* wells_out) {
  PMX_TRY({
    PMX_CHECK(h && p && out7, "null argument");
    bind(h);
    h->m->member_convergence(p, out7, wells_out);
  })
}

int pmx_ow_mask_residual(pmx_ow* h, const double* mask) {
  PMX_TRY({
    PMX_CHECK(h && mask, "null argument");
    bind(h);
    h->m->mask_residual(mask);
  })
}

// ---------------------------------------------------------------------------
// three-phase black-oil model
// ---------------------------------------------------------------------------
int pmx_bo_create(int32_t device, void* hip_stream, const double* scalars, int32_t n_pvdg, const double* pvdg,
                  int32_t n_key, const double* keys, const int32_t* branch_ptr, const double* branch,
                  int32_t n_swof, const double* swof, int32_t n_sgof, const double* sgof, pmx_bo** out) {
  PMX_TRY({
    PMX_CHECK(out && scalars && pvdg && keys && branch_ptr && branch && swof && sgof, "null argument");
    auto h = std::make_unique<pmx_bo>();
    h->device = device;
    bind(h.get());
    h->m = std::make_unique<BOModelDev>((hipStream_t)hip_stream);
    h->m->fluid.build(scalars, n_pvdg, pvdg, n_key, keys, branch_ptr, branch, n_swof, swof, n_sgof, sgof);
    h->m->fluid.upload(h->m->stream);
    PMX_HIP(hipStreamSynchronize(h->m->stream));
    pmx_solver_defaults(&h->last_opts);
    *out = h.release();
  })
}

int pmx_bo_destroy(pmx_bo* h) {
  PMX_TRY({
    if (h) {
      bind(h);
      delete h;
    }
  })
}

int pmx_bo_set_topology(pmx_bo* h, int32_t nc, int32_t nconn, const int32_t* N, const double* T, const double* gdz,
                        const double* pv) {
  PMX_TRY({
    PMX_CHECK(h && N && T && gdz && pv, "null argument");
    bind(h);
    h->m->set_topology(nc, nconn, N, T, gdz, pv);
  })
}

int pmx_bo_set_wells(pmx_bo* h, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells, const double* WI,
                     const int32_t* type, const double* val, const double* compi) {
  PMX_TRY({
    PMX_CHECK(h && (nw == 0 || (perf_ptr && perf_cells && WI && type && val && compi)), "null argument");
    bind(h);
    for (int w = 0; w < nw; ++w) PMX_CHECK(type[w] == 0 || type[w] == 1, "unsupported well control");
    h->m->set_wells(nw, perf_ptr, perf_cells, WI, type, val, compi);
  })
}

int pmx_bo_set_lines(pmx_bo* h, int32_t nlines, const int32_t* line_ptr, const int32_t* cells) {
  PMX_TRY({
    PMX_CHECK(h && line_ptr && cells && nlines >= 0, "null argument");
    cpr_set_lines(h->m->cpr, nlines, line_ptr, cells);
  })
}

int pmx_bo_begin_step(pmx_bo* h, const double* p0, const double* s0, const double* rs0, const int32_t* status0) {
  PMX_TRY({
    PMX_CHECK(h && p0 && s0 && rs0 && status0, "null argument");
    bind(h);
    h->m->begin_step(p0, s0, rs0, status0);
  })
}

int pmx_bo_assemble(pmx_bo* h, const double* p, const double* s, const double* rs, const int32_t* status,
                    const double* q, const double* bhp, double dt) {
  PMX_TRY({
    PMX_CHECK(h && p && s && rs && status, "null argument");
    PMX_CHECK(h->m->nw == 0 || (q && bhp), "well state required");
    PMX_CHECK(dt > 0, "dt must be positive");
    bind(h);
    h->m->assemble(p, s, rs, status, q, bhp, dt);
  })
}

int pmx_bo_residual(pmx_bo* h, double* out) {
  PMX_TRY({
    PMX_CHECK(h && out, "null argument");
    bind(h);
    PMX_CHECK(h->m->r.n >= h->m->S.n(), "assemble first");
    PMX_HIP(hipMemcpyAsync(out, h->m->r.p, sizeof(double) * h->m->S.n(), hipMemcpyDeviceToDevice, h->m->stream));
  })
}

int pmx_bo_convergence(pmx_bo* h, const double* p, const double* rs, const int32_t* status, double* out12,
                       double* wells_out) {
  PMX_TRY({
    PMX_CHECK(h && p && rs && status && out12, "null argument");
    bind(h);
    h->m->convergence(p, rs, status, out12, wells_out);
  })
}

int pmx_bo_solve(pmx_bo* h, const pmx_solver_options* o, double pmean, double rsmean, double* dx,
                 pmx_solve_info* info) {
  try {
    PMX_CHECK(h && dx, "null argument");
    bind(h);
    SolveInfo si;
    LinearSolverOptions lo = to_opts(o);
    if (o) h->last_opts = *o;
    const int st = h->m->solve(lo, dx, si, pmean, rsmean);
    fill_info(info, si);
    if (st < 0) {
      g_err = "linear solve produced a non-finite residual";
      return -2;
    }
    return st;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int pmx_bo_update(pmx_bo* h, const double* dx, double ds_max, double* p, double* s, double* rs, int32_t* status,
                  double* q, double* bhp) {
  PMX_TRY({
    PMX_CHECK(h && dx && p && s && rs && status, "null argument");
    PMX_CHECK(h->m->nw == 0 || (q && bhp), "well state required");
    bind(h);
    h->m->update(dx, ds_max, p, s, rs, status, q, bhp);
  })
}

int pmx_bo_set_controls(pmx_bo* h, const double* coef, const double* val, const int32_t* status,
                        const double* rscale) {
  PMX_TRY({
    PMX_CHECK(h && rscale, "null argument");
    PMX_CHECK(h->m->nw == 0 || (coef && val && status), "well controls required");
    for (int k = 0; k < h->m->nw; ++k) PMX_CHECK(status[k] == 0 || status[k] == 1, "well status is 0 or 1");
    bind(h);
    h->m->set_controls(coef, val, status, rscale);
  })
}

int pmx_bo_set_connection_drops(pmx_bo* h, const double* cdp) {
  PMX_TRY({
    PMX_CHECK(h && (h->m->nw == 0 || cdp), "null argument");
    bind(h);
    h->m->set_connection_drops(cdp);
  })
}

int pmx_bo_perf_volume_rates(pmx_bo* h, double* out) {
  PMX_TRY({
    PMX_CHECK(h && (h->m->nw == 0 || out), "null argument");
    bind(h);
    h->m->perf_volume_rates(out);
  })
}

int pmx_bo_set_update_rule(pmx_bo* h, int32_t rule, double dp_max_rel, double drs_max) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    PMX_CHECK(dp_max_rel >= 0.0 && drs_max >= 0.0, "update limits must be non-negative");
    h->m->set_update_rule(rule, dp_max_rel, drs_max);
  })
}

int pmx_bo_cell_densities(pmx_bo* h, const double* p, const double* s, const double* rs, const int32_t* status,
                          double* out) {
  PMX_TRY({
    PMX_CHECK(h && p && s && rs && status && out, "null argument");
    bind(h);
    h->m->cell_densities(p, s, rs, status, out);
  })
}

int pmx_bo_matvec(pmx_bo* h, const double* x, double* y) {
  PMX_TRY({
    PMX_CHECK(h && x && y, "null argument");
    bind(h);
    DArr<double> hs;
    hs.alloc((size_t)std::max(h->m->nw, 1) * kMaxMix);
    system_matvec(h->m->S, x, y, hs.p, h->m->stream);
    PMX_HIP(hipStreamSynchronize(h->m->stream));
  })
}

int pmx_bo_nnzb(pmx_bo* h, int32_t* nnzb) {
  PMX_TRY({
    PMX_CHECK(h && nnzb, "null argument");
    *nnzb = h->m->S.nnzb;
  })
}

int pmx_bo_export_pattern(pmx_bo* h, int32_t* rowptr, int32_t* col) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    bind(h);
    auto& S = h->m->S;
    if (rowptr)
      PMX_HIP(hipMemcpyAsync(rowptr, S.rowptr.p, sizeof(int) * (S.nc + 1), hipMemcpyDeviceToHost, h->m->stream));
    if (col) PMX_HIP(hipMemcpyAsync(col, S.col.p, sizeof(int) * S.nnzb, hipMemcpyDeviceToHost, h->m->stream));
    PMX_HIP(hipStreamSynchronize(h->m->stream));
  })
}

int pmx_bo_export_blocks(pmx_bo* h, double* vals) {
  PMX_TRY({
    PMX_CHECK(h && vals, "null argument");
    bind(h);
    auto& S = h->m->S;
    PMX_HIP(hipMemcpyAsync(vals, S.val.p, sizeof(double) * 9 * S.nnzb, hipMemcpyDeviceToDevice, h->m->stream));
  })
}

int pmx_bo_profile(pmx_bo* h, int32_t enable) {
  PMX_TRY({
    PMX_CHECK(h, "null argument");
    cpr_profile(h->m->cpr, enable != 0);
  })
}

int pmx_bo_launch_shapes(pmx_bo* h, int32_t* out, int32_t cap, int32_t* count) {
  PMX_TRY({
    PMX_CHECK(h && out && count && cap >= 2, "bad argument");
    *count = cpr_launch_shapes(h->m->cpr, h->m->S, out, cap);
  })
}

int pmx_bo_profile_read(pmx_bo* h, double* out4) {
  PMX_TRY({
    PMX_CHECK(h && out4, "null argument");
    bind(h);
    cpr_profile_read(h->m->cpr, out4);
  })
}


int pmx_sys_create(int32_t device, void* hip_stream, int32_t B, int32_t nc, const int32_t* rowptr,
                   const int32_t* col, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells,
                   pmx_sys** out) {
  try {
    PMX_CHECK(out != nullptr && rowptr != nullptr && col != nullptr, "null argument");
    PMX_CHECK(B == 2 || B == 3, "block size must be 2 (oil-water) or 3 (black-oil)");
    PMX_CHECK(nc > 0 && nw >= 0, "empty system");
    PMX_CHECK(nw == 0 || (perf_ptr != nullptr && perf_cells != nullptr), "wells need their perforations");
    auto h = std::make_unique<pmx_sys>();
    h->device = device;
    PMX_HIP(hipSetDevice(device));
    h->stream = (hipStream_t)hip_stream;
    h->S.B = B;
    h->S.M = B + 1;
    h->S.X = B;
    HostPattern pat;
    pat.n = nc;
    pat.rowptr.assign(rowptr, rowptr + nc + 1);
    PMX_CHECK(pat.rowptr[0] == 0, "rowptr must start at 0");
    const int nnzb = pat.rowptr[nc];
    pat.col.assign(col, col + nnzb);
    pat.diag.assign(nc, -1);
    for (int r = 0; r < nc; ++r) {
      PMX_CHECK(pat.rowptr[r + 1] >= pat.rowptr[r], "rowptr must not decrease");
      for (int k = pat.rowptr[r]; k < pat.rowptr[r + 1]; ++k) {
        PMX_CHECK(pat.col[k] >= 0 && pat.col[k] < nc, "block column out of range");
        PMX_CHECK(k == pat.rowptr[r] || pat.col[k] > pat.col[k - 1], "block columns must ascend within a row");
        if (pat.col[k] == r) pat.diag[r] = k;
      }
      PMX_CHECK(pat.diag[r] >= 0, "every block row needs its diagonal block");
    }
    h->S.set_pattern(pat, h->stream);
    std::vector<int> pp(nw + 1, 0), pc;
    if (nw > 0) {
      pp.assign(perf_ptr, perf_ptr + nw + 1);
      PMX_CHECK(pp[0] == 0, "perf_ptr must start at 0");
      for (int w = 0; w < nw; ++w) PMX_CHECK(pp[w + 1] > pp[w], "every well needs a perforation");
      pc.assign(perf_cells, perf_cells + pp[nw]);
    }
    h->S.set_wells(nw, pp, pc, h->stream);
    h->S.has_mix = false;
    h->h_rowptr = pat.rowptr;
    h->h_col = pat.col;
    h->cpr = cpr_create();
    h->w.alloc(B);
    PMX_HIP(hipStreamSynchronize(h->stream));
    *out = h.release();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int pmx_sys_destroy(pmx_sys* h) {
  try {
    if (h) {
      PMX_HIP(hipSetDevice(h->device));
      delete h;
    }
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int pmx_sys_set_values(pmx_sys* h, const double* val, const double* Bp, const double* Cp, const double* Dp) {
  try {
    PMX_CHECK(h != nullptr && val != nullptr, "null argument");
    PMX_HIP(hipSetDevice(h->device));
    BlockSystem& S = h->S;
    const size_t B = S.B, M = S.M;
    PMX_HIP(hipMemcpyAsync(S.val.p, val, sizeof(double) * S.nnzb * B * B, hipMemcpyDeviceToDevice, h->stream));
    if (S.nw > 0) {
      PMX_CHECK(Bp != nullptr && Cp != nullptr && Dp != nullptr, "a system with wells needs Bp, Cp and Dp");
      PMX_HIP(hipMemcpyAsync(S.Bp.p, Bp, sizeof(double) * S.nperf * B * M, hipMemcpyDeviceToDevice, h->stream));
      PMX_HIP(hipMemcpyAsync(S.Cp.p, Cp, sizeof(double) * S.nperf * M * B, hipMemcpyDeviceToDevice, h->stream));
      PMX_HIP(hipMemcpyAsync(S.Dp.p, Dp, sizeof(double) * S.nw * M * M, hipMemcpyDeviceToDevice, h->stream));
    }
    if (!h->have_values) {
      // static aggregation from the first values: |a_pp| of each block
      std::vector<double> pp(S.nnzb);
      std::vector<double> blk((size_t)S.nnzb * B * B);
      PMX_HIP(hipMemcpyAsync(blk.data(), S.val.p, sizeof(double) * blk.size(), hipMemcpyDeviceToHost, h->stream));
      PMX_HIP(hipStreamSynchronize(h->stream));
      for (int k = 0; k < S.nnzb; ++k) pp[k] = std::fabs(blk[(size_t)k * B * B]);
      LinearSolverOptions o;
      cpr_build_hierarchy(h->cpr, S, pp, o, h->stream);
      h->have_values = true;
    }
    PMX_HIP(hipStreamSynchronize(h->stream));
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int pmx_sys_matvec(pmx_sys* h, const double* x, double* y) {
  try {
    PMX_CHECK(h != nullptr && x != nullptr && y != nullptr, "null argument");
    PMX_HIP(hipSetDevice(h->device));
    DArr<double> scratch;
    scratch.alloc(std::max<size_t>((size_t)h->S.nw * h->S.M, 1));
    system_matvec(h->S, x, y, scratch.p, h->stream);
    PMX_HIP(hipStreamSynchronize(h->stream));
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return 0;
}

int pmx_sys_set_pressure_solver(pmx_sys* h, pmx_pressure_solver fn, void* user) {
  PMX_TRY({
    PMX_CHECK(h != nullptr, "null argument");
    cpr_set_pressure_solver(h->cpr, fn, user);
  })
}

int pmx_sys_solve(pmx_sys* h, const pmx_solver_options* opts, const double* w, const double* rhs, double* x,
                  pmx_solve_info* info) {
  int status = 0;
  try {
    PMX_CHECK(h != nullptr && w != nullptr && rhs != nullptr && x != nullptr, "null argument");
    PMX_CHECK(h->have_values, "set the values before solving");
    PMX_HIP(hipSetDevice(h->device));
    LinearSolverOptions o = to_opts(opts);
    PMX_HIP(hipMemcpyAsync(h->w.p, w, sizeof(double) * h->S.B, hipMemcpyHostToDevice, h->stream));
    auto t0 = std::chrono::steady_clock::now();
    cpr_setup(h->cpr, h->S, h->w.p, o, h->stream);
    PMX_HIP(hipStreamSynchronize(h->stream));
    const double setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    SolveInfo si;
    status = cpr_solve(h->cpr, h->S, rhs, x, o, si, h->stream);
    si.setup_ms = setup_ms;
    fill_info(info, si);
    if (status < 0) {
      g_err = "linear solve produced a non-finite residual";
      return -2;
    }
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
  return status;
}


int pmx_cp_process(int32_t nx, int32_t ny, int32_t nz, const double* coord, const double* zcorn,
                   const int32_t* actnum, pmx_cp** out) {
  PMX_TRY({
    PMX_CHECK(out, "null argument");
    *out = nullptr;
    std::unique_ptr<pmx_cp> h(new pmx_cp());
    cp_process(nx, ny, nz, coord, zcorn, actnum, h->g);
    *out = h.release();
  })
}

int pmx_cp_destroy(pmx_cp* h) {
  delete h;
  return 0;
}

int pmx_cp_sizes(pmx_cp* h, int32_t* out4) {
  PMX_TRY({
    PMX_CHECK(h && out4, "null argument");
    out4[0] = h->g.nc;
    out4[1] = h->g.nf;
    out4[2] = h->g.nn;
    out4[3] = (int32_t)h->g.face_nodes.size();
  })
}

int pmx_cp_topology(pmx_cp* h, int32_t* index_map, int32_t* neighbors, int32_t* tag, int32_t* node_pos,
                    int32_t* face_nodes, double* coords, int32_t* face_pos, int32_t* cell_faces, double* cell_dims) {
  PMX_TRY({
    PMX_CHECK(h && index_map && neighbors && tag && node_pos && face_nodes && coords && face_pos && cell_faces &&
                  cell_dims,
              "null argument");
    const CpGrid& g = h->g;
    std::copy(g.index_map.begin(), g.index_map.end(), index_map);
    std::copy(g.nb.begin(), g.nb.end(), neighbors);
    std::copy(g.tag.begin(), g.tag.end(), tag);
    std::copy(g.node_pos.begin(), g.node_pos.end(), node_pos);
    std::copy(g.face_nodes.begin(), g.face_nodes.end(), face_nodes);
    std::copy(g.coords.begin(), g.coords.end(), coords);
    std::copy(g.face_pos.begin(), g.face_pos.end(), face_pos);
    std::copy(g.cell_faces.begin(), g.cell_faces.end(), cell_faces);
    std::copy(g.cell_dims.begin(), g.cell_dims.end(), cell_dims);
  })
}

int pmx_cp_geometry(pmx_cp* h, int32_t device, void* hip_stream, double* face_normals, double* face_areas,
                    double* face_centroids, double* cell_volumes, double* cell_centroids) {
  PMX_TRY({
    PMX_CHECK(h && face_normals && face_areas && face_centroids && cell_volumes && cell_centroids, "null argument");
    PMX_HIP(hipSetDevice(device));
    cp_geometry(h->g, (hipStream_t)hip_stream, face_normals, face_areas, face_centroids, cell_volumes,
                cell_centroids);
  })
}

}  // extern "C"
