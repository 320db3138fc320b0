/*
 * pmx -- C ABI of the MI355X-native TPFA Newton step.
 *
 * Every array argument documented "device" is a caller-owned device pointer
 * (e.g. a torch tensor's data_ptr()); "host" arrays are read during the call
 * only.  One handle is bound to one HIP stream and must not be used from two
 * host threads at once; every call is ordered on that stream.  Functions
 * return 0 on success, 1 when an iterative solve did not converge, and a
 * negative value on error (message from pmx_last_error()).
 *
 * Vector layouts on the device:
 *   cell unknowns   cell-major interleaved: (p_0, sW_0, p_1, sW_1, ...)
 *   well unknowns   well-major after the cells: (qWs_0, qOs_0, bhp_0, ...)
 *   residual rows   (water_0, oil_0, ...), then (eq_w, eq_o, ctrl) per well
 * The reference orders the same unknowns variable-blocked,
 * [p(nc), sW(nc), qWs(nw), qOs(nw), bhp(nw)]; the Python layer permutes.
 */
#ifndef PMX_H
#define PMX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmx_ow pmx_ow;

typedef struct pmx_solver_options {
  double rtol;              /* ||b - J x|| <= rtol ||b|| */
  int32_t restart;
  int32_t maxiter;          /* restart cycles, as scipy's gmres counts them */
  int32_t precision;        /* 64: FP64 preconditioner, 32: FP32 preconditioner */
  int32_t use_pressure_stage;
  int32_t pre_smooth, post_smooth;
  double jacobi_omega;
  double coarse_alpha;
  double theta;             /* aggregation strength threshold */
  int32_t coarse_size;
  int32_t pairwise_passes;
  int32_t reorthogonalize;
  int32_t verbose;
  int32_t local_stage;      /* 0: cell block-Jacobi, 1: coloured line Gauss-Seidel, 2: block ILU(0),
                               level-scheduled (MRST's ILU(0), gmres_ilu_solver_ad.py:119-209) */
  int32_t pressure_cycles;  /* V-cycles per pressure-stage application (stationary iteration) */
  int32_t local_sweeps;     /* Gauss-Seidel sweeps over the line colours */
  int32_t use_graphs;       /* 1: replay the preconditioner as a captured HIP graph */
  int32_t coarse_block_rows; /* AMG levels up to this many rows run in one workgroup */
  int32_t line_segments;    /* 1: lines of 32+ cells solved by segments (spikes and a Schur system on
                               separator cells); 0: each line by one twisted walk (default) */
  int32_t autotune;         /* 1: choose launch shapes by timing (each AMG level's sweep layout, lanes per
                               row and folded launches; the SpMV's lanes per block row) at the first
                               setup of each size class, kept in a process-wide table; 0: fixed shapes */
  int32_t level_shape;      /* -1 (default): shapes as above; 0-6: every AMG level above the coarse cycle
                               takes this shape (0 sweeps from the ELL copy where the level has one,
                               1-3 CSR with 4, 8, 16 lanes per row, 4-6 the same with folded launches) */
  int32_t dot_shape;        /* 0 (default): the projections' shape as above; 2, 4, 8 (elements per thread,
                               shuffle sums) or 104, 108 (4 or 8, LDS sums): that shape */
} pmx_solver_options;

typedef struct pmx_solve_info {
  int32_t iterations, restarts;
  double rel_residual, setup_ms, solve_ms;
  int32_t levels, coarse_n;
} pmx_solve_info;

/* Last error message of the calling thread. */
const char* pmx_last_error(void);
/* Default solver options (rtol 1e-8, restart 50, maxiter 200: the reference's
 * CPRSolver defaults, pymrst/modules/ad_core/solvers/linear_solvers.py:113-120). */
void pmx_solver_defaults(pmx_solver_options* o);

/* ---- oil-water model (TwoPhaseOilWaterModelAD,
 *      pymrst/modules/ad_blackoil/models/two_phase_oil_water_model.py:79-657) */

/* fluid[10] = {mu_w, mu_o, rhoS_w, rhoS_o, n_w, n_o, c_w, c_o, p_ref, c_r};
 * c_r < 0 means no rock compressibility (the legacy model's constant pv). */
int pmx_ow_create(int32_t device, void* hip_stream, const double* fluid, pmx_ow** out);
int pmx_ow_destroy(pmx_ow* h);

/* Topology: nc cells, nconn internal connections N[2*nconn] (host int32,
 * setup_operators_tpfa.py:98-110), T[nconn] transmissibilities, gdz[nconn] =
 * g (z[N1] - z[N0]) (two_phase_oil_water_model.py:157-162), pv[nc] pore
 * volumes.  All host arrays. */
int pmx_ow_set_topology(pmx_ow* h, int32_t nc, int32_t nconn, const int32_t* N, const double* T,
                        const double* gdz, const double* pv);

/* Wells (add_well, pymrst/params/wells_and_bc/add_well.py:16-126): perf_ptr[nw+1],
 * perf_cells[nperf], WI[nperf], type[nw] (0 bhp, 1 rate), val[nw],
 * compi[2*nw].  Host arrays. */
int pmx_ow_set_wells(pmx_ow* h, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells,
                     const double* WI, const int32_t* type, const double* val, const double* compi);

/* Residual and Jacobian at (p, sW, qWs, qOs, bhp) with the previous state
 * (p0, sW0), time step dt and per-connection cdp[nperf] (device; may be NULL
 * for zeros).  All state arrays device.  Replaces assemble + combine_equations
 * (two_phase_oil_water_model.py:493-546, utils/adi/operators.py:204-211). */
int pmx_ow_assemble(pmx_ow* h, const double* p, const double* sw, const double* qws, const double* qos,
                    const double* bhp, const double* p0, const double* sw0, double dt, const double* cdp);

/* Copy the last residual (2*nc + 3*nw, device layout) into out (device). */
int pmx_ow_residual(pmx_ow* h, double* out);

/* Convergence reductions of the last residual (_converged,
 * two_phase_oil_water_model.py:549-567): out7 (host) = {max|r_w/pv|,
 * max|r_o/pv|, sum r_w, sum r_o, sum 1/b_w(p), sum 1/b_o(p), sum p};
 * wells_out (host, 3*nw) = well rows of the residual. */
int pmx_ow_convergence(pmx_ow* h, const double* p, double* out7, double* wells_out);

/* The facility's controls (FacilityModel._control_equation,
 * pymrst/modules/ad_core/models/facilities/facility_model.py:2589-2658, and
 * _single_well :2434-2549): each well's control row is the linear form
 * coef[3w] qWs + coef[3w+1] qOs + coef[3w+2] bhp - val[w] (a bhp target,
 * a total, single-phase or liquid rate, a zero target's total rate, a
 * reservoir-rate form with control densities); status[w] = 0 shuts a well
 * (rows qWs = qOs = 0 and bhp = the mean of its perforations' pressures);
 * the two rate-closure rows of every open well are scaled by rscale[2]
 * (MRST's rhoScale).  Host arrays; applies from the next pmx_ow_assemble.
 * pmx_ow_set_wells resets them to its bhp / rate controls, unscaled. */
int pmx_ow_set_controls(pmx_ow* h, const double* coef, const double* val, const int32_t* status,
                        const double* rscale);

/* pmx_ow_convergence with the current pore volume pv (1 + c_r (p - p_ref))
 * in the maxima (GenericBlackOilModelAD._convergence_scaling,
 * pymrst/modules/ad_blackoil/models/generic_black_oil_model.py:604-638) and
 * its sum as out8[7]. */
int pmx_ow_convergence_current_pv(pmx_ow* h, const double* p, double* out8, double* wells_out);

/* Reservoir volume rates of every perforation at the last pmx_ow_assemble
 * (host, 2*nperf, [water, oil] per perforation: a producing perforation's
 * phase rates, an injecting one's total rate times compi), the facility's
 * cq_vol (facility_model.py:2512-2528) that a converged step hands to the
 * next one's wellbore column (generic_black_oil_model.py:569-601). */
int pmx_ow_perf_volume_rates(pmx_ow* h, double* out);

/* pmx_ow_convergence without any allocation: the same reductions and
 * outputs, with the partial sums kept in a static per-GPU array of the
 * library (512 rows of 8 doubles plus the 8 results, 32,832 bytes, held
 * until the process ends) instead of handle memory.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two streams, two host
 * threads) race.  Any cell count fits; a reduction that would need more
 * partial rows than the array holds is rejected with an error. */
int pmx_ow_convergence_noalloc(pmx_ow* h, const double* p, double* out7, double* wells_out);

/* Solve J dx = -r (CPR-FGMRES) into dx (device, n = 2*nc + 3*nw); pmean is
 * the mean pressure for the quasi-IMPES weights.  Replaces
 * _solve_increment / CPRSolver.solve (linear_solvers.py:426-545). */
int pmx_ow_solve(pmx_ow* h, const pmx_solver_options* o, double pmean, double* dx, pmx_solve_info* info);

/* Apply x += sign*dx with the Newton update limits (step,
 * two_phase_oil_water_model.py:629-651).  All device. */
int pmx_ow_update(pmx_ow* h, const double* dx, double sign, double ds_max, double dp_max_rel, double* p, double* sw,
                  double* qws, double* qos, double* bhp);

/* y = J x for the last Jacobian (device vectors of length 2*nc + 3*nw). */
int pmx_ow_matvec(pmx_ow* h, const double* x, double* y);

/* Export the last Jacobian: block pattern (host rowptr[nc+1], col[nnzb]) and
 * values (device, nnzb*4 row-major 2x2 blocks).  Either pointer may be NULL. */
int pmx_ow_nnzb(pmx_ow* h, int32_t* nnzb);
int pmx_ow_export_pattern(pmx_ow* h, int32_t* rowptr, int32_t* col);
int pmx_ow_export_blocks(pmx_ow* h, double* vals);

/* Connection surface rates (cq_w, cq_o) per perforation of the last
 * assembly, into host out[2*nperf]. */
int pmx_ow_connection_rates(pmx_ow* h, double* out);

/* Lines of cells for the line local stage of CPR (host arrays:
 * line_ptr[nlines+1], cells[line_ptr[nlines]]; consecutive cells of a line
 * must be connected).  Without this call lines are detected from the
 * transmissibilities.  No reference counterpart: the reference's local
 * stages are block-Jacobi and ILU (linear_solvers.py:292-363). */
int pmx_ow_set_lines(pmx_ow* h, int32_t nlines, const int32_t* line_ptr, const int32_t* cells);

/* Kernel timing with HIP events on the handle's stream: enable (1) / disable
 * (0, also clears); out4 (host) = {BCSR SpMV launches inside FGMRES, their
 * total ms, preconditioner applications, their total ms}. */
int pmx_ow_profile(pmx_ow* h, int32_t enable);
int pmx_ow_profile_read(pmx_ow* h, double* out4);
/* The line Gauss-Seidel solve kernel timed launch by launch with HIP events
 * while profiling is on, in preconditioner applications that run without
 * graph replay (pmx_ow_precondition with use_graphs = 0); out3 (host) =
 * {launches timed, their total ms, their algorithmic bytes}.  No counterpart
 * in the reference (measurement only). */
int pmx_ow_profile_line_gs(pmx_ow* h, double* out3);
/* The launch shapes in use after a solve (pmx_solver_options.autotune):
 * out[0] = the Jacobian SpMV's lanes per block row (0 before the first
 * product), out[1] = number of AMG levels reported, then per level above
 * the single-workgroup coarse cycle {rows, sweeps from the ELL copy (1) or
 * CSR (0), lanes per CSR row, folded launches (1/0), tuned (1/0)}, then the
 * Gram-Schmidt projections' shape (2, 4, 8: elements per thread with
 * shuffle sums; 104, 108: 4 or 8 with LDS sums; 0 before the first solve);
 * *count = ints written (at most cap).  No counterpart in the reference
 * (launch configuration of this implementation). */
int pmx_ow_launch_shapes(pmx_ow* h, int32_t* out, int32_t cap, int32_t* count);
/* The process-wide autotuning table (size class -> chosen shape): number of
 * entries / empty it, so that the next setups time their candidates again.
 * No counterpart in the reference. */
int pmx_autotune_entries(int32_t* count);
int pmx_autotune_reset(void);
/* The table as a text file (one entry per line: launch kind, element size,
 * rows class, entries-per-row class, extra, choice): written, or read back
 * so that another process uses the same shapes without timing them (a run
 * under a profiler, whose per-dispatch overhead would bias the timings).
 * *count (may be NULL) = entries written / read. */
int pmx_autotune_save(const char* path, int32_t* count);
int pmx_autotune_load(const char* path, int32_t* count);

/* One preconditioner application z = M^{-1} r after a solve's setup (device). */
int pmx_ow_precondition(pmx_ow* h, const pmx_solver_options* o, const double* r, double* z);

/* Independent members (an ensemble solved as one system): member s owns cells
 * [cell_ptr[s], cell_ptr[s+1]) and wells [well_ptr[s], well_ptr[s+1]) (host
 * arrays of nseg+1).  FGMRES then runs one Krylov recurrence per member
 * (own dots, Hessenberg, Givens, convergence and restarts) while the
 * preconditioner and the SpMV cover all members in each launch.  No
 * reference counterpart: the reference loops over realisations. */
int pmx_ow_set_segments(pmx_ow* h, int32_t nseg, const int32_t* cell_ptr, const int32_t* well_ptr);
/* Multiply the assembled residual by mask (device, one value per unknown):
 * zero rows of members that have converged, so their Newton update is zero
 * while the others iterate.  No reference counterpart. */
int pmx_ow_mask_residual(pmx_ow* h, const double* mask);
/* Members of a union model (an ensemble): member m owns the cells
 * [cell_ptr[m], cell_ptr[m+1]) (host array of nmem+1, ascending, from 0 to
 * the model's cell count).  Set once per model; pmx_ow_member_convergence
 * then reduces each member's cells on its own.  No reference counterpart:
 * the reference loops over realisations. */
int pmx_ow_set_members(pmx_ow* h, int32_t nmem, const int32_t* cell_ptr);
/* pmx_ow_convergence per member: the quantities of the reference's
 * _converged (two_phase_oil_water_model.py:549-567) for each member's own
 * cells, out7 (host, nmem rows of 7) = {max|r_w/pv|, max|r_o/pv|, sum r_w,
 * sum r_o, sum 1/b_w(p), sum 1/b_o(p), sum p}, reduced in a fixed order;
 * wells_out (host, 3 nw, or NULL) = the well residual rows. */
int pmx_ow_member_convergence(pmx_ow* h, const double* p, double* out7, double* wells_out);
/* Per-cell step lengths for the accumulation terms of the following
 * pmx_ow_assemble calls (device array of nc values, caller-owned, read at
 * every assembly), so that ensemble members advance with their own time
 * steps; NULL returns to the scalar dt.  No reference counterpart. */
int pmx_ow_set_cell_dt(pmx_ow* h, const double* dt_cell);

/* ---- discrete adjoint (pymrst/modules/optimization/adjoint.py:73-225) */
/* Per connection f (host, 2*nconn): T_f/T_h at N0 and at N1, the derivative
 * of the face transmissibility with respect to the log-permeability of each
 * side, divided by T_f (dT_f/dln k_c = T_f^2/T_h,c; solvers/compute_trans.py
 * and get_face_transmissibility.py give T_h and T_f). */
int pmx_ow_set_trans_fractions(pmx_ow* h, const double* frac);
/* (dR/dx)^T lam = rhs for the Jacobian of the last pmx_ow_assemble
 * (adjoint.py:203 `spsolve(a.T, rhs)`), by CPR-FGMRES on the transposed
 * system; rhs and lam are device vectors in the device unknown order.
 * Returns 0 converged, 1 not converged, <0 error. */
int pmx_ow_adjoint_solve(pmx_ow* h, const pmx_solver_options* o, double pmean, const double* rhs, double* lam,
                         pmx_solve_info* info);
/* out = (dR_n/dx_{n-1})^T lam (adjoint.py:85-107, accumulation only; zero
 * on the well unknowns); p0, s0 are the previous state (device). */
int pmx_ow_prev_state_transpose(pmx_ow* h, const double* p0, const double* s0, double dt, const double* lam,
                                double* out);
/* grad[c] += (dR/dln k_c)^T lam at the state (p, s, qws, qos, bhp) with the
 * connection head cdp (device; nullptr: the handle's), through the face
 * transmissibilities and the Peaceman well indices. */
int pmx_ow_dlogk(pmx_ow* h, const double* p, const double* s, const double* qws, const double* qos,
                 const double* bhp, const double* cdp, const double* lam, double* grad);

/* ---- three-phase live-oil black-oil model (ThreePhaseBlackOilModelAD,
 *      pymrst/modules/ad_blackoil/models/three_phase_black_oil_model.py).
 * Device unknowns: cell-major (p, sW, x) with x = rs where status == 1 and
 * x = sG where status == 3, then well-major (qWs, qOs, qGs, bhp); residual
 * rows (water, oil, gas) per cell and (eq_w, eq_o, eq_g, ctrl) per well;
 * 3x3 row-major Jacobian blocks.  State arrays are device memory: p[nc],
 * s[3*nc] (sW, sO, sG per cell), rs[nc], status[nc] (int32: 1 or 3),
 * q[3*nw] (qWs, qOs, qGs per well), bhp[nw]. */
typedef struct pmx_bo pmx_bo;

/* Create the model with the deck fluid (init_deck_blackoil_fluid,
 * pymrst/modules/ad_props/props/init_deck_adi_fluid.py:63-194), host arrays:
 * scalars[9] = {rhoS_w, rhoS_o, rhoS_g (DENSITY), PVTW p_ref, B_ref, c, mu,
 * ROCK p_ref, c}; pvdg[3*n_pvdg] rows (p, Bg, mu_g); PVTO as n_key Rs keys
 * with rows branch[3*i] = (p, Bo, mu_o) for i in [branch_ptr[k],
 * branch_ptr[k+1]) (first row saturated); swof[4*n_swof] and sgof[4*n_sgof]
 * rows (s, kr, kro, pc). */
int pmx_bo_create(int32_t device, void* hip_stream, const double* scalars, int32_t n_pvdg, const double* pvdg,
                  int32_t n_key, const double* keys, const int32_t* branch_ptr, const double* branch,
                  int32_t n_swof, const double* swof, int32_t n_sgof, const double* sgof, pmx_bo** out);
int pmx_bo_destroy(pmx_bo* h);
/* Connections, transmissibilities, gravity head g*dz per connection and
 * reference pore volumes (host), as the model's constructor builds them
 * (three_phase_black_oil_model.py:116-160). */
int pmx_bo_set_topology(pmx_bo* h, int32_t nc, int32_t nconn, const int32_t* N, const double* T, const double* gdz,
                        const double* pv);
/* Wells (host): perforations grouped by well, WI per perforation, type 0 bhp /
 * 1 total surface rate, target, compi[3*nw] (water, oil, gas) injection mix. */
int pmx_bo_set_wells(pmx_bo* h, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells, const double* WI,
                     const int32_t* type, const double* val, const double* compi);
/* Lines for the line local stage of CPR (as pmx_ow_set_lines). */
int pmx_bo_set_lines(pmx_bo* h, int32_t nlines, const int32_t* line_ptr, const int32_t* cells);
/* Record the start-of-step accumulation terms of state0 (device). */
int pmx_bo_begin_step(pmx_bo* h, const double* p0, const double* s0, const double* rs0, const int32_t* status0);
/* Residual and Jacobian at the state (assemble :314-381 with _well
 * :383-436); device arrays. */
int pmx_bo_assemble(pmx_bo* h, const double* p, const double* s, const double* rs, const int32_t* status,
                    const double* q, const double* bhp, double dt);
/* Copy the last residual into out (device, 3*nc + 4*nw). */
int pmx_bo_residual(pmx_bo* h, double* out);
/* Reductions for _converged (:512-545), host out12 = {max|r_a/pv(p)| (3),
 * sum r_a (3), sum 1/b_a (3), sum p, sum rs, sum pv(p)}; wells_out (host,
 * 4*nw) gets the well rows.  p, rs, status are the device state. */
int pmx_bo_convergence(pmx_bo* h, const double* p, const double* rs, const int32_t* status, double* out12,
                       double* wells_out);
/* Newton increment J dx = -r (step :547-571, a direct solve there) by
 * CPR-FGMRES with pressure weights at (pmean, rsmean). */
int pmx_bo_solve(pmx_bo* h, const pmx_solver_options* o, double pmean, double rsmean, double* dx,
                 pmx_solve_info* info);
/* State update with the variable switch and the flash (update_state
 * :439-470, _flash :472-509), in place on the device state. */
int pmx_bo_update(pmx_bo* h, const double* dx, double ds_max, double* p, double* s, double* rs, int32_t* status,
                  double* q, double* bhp);
/* The facility's well equations for GenericBlackOilModelAD
 * (FacilityModel.well_equations / _single_well / _control_equation,
 * pymrst/modules/ad_core/models/facilities/facility_model.py:2197-2292,
 * 2434-2549, 2589-2658), held until changed, all host arrays: each well's
 * control row c_w qWs + c_o qOs + c_g qGs + c_b bhp - val (coef[4*nw],
 * val[nw]; the hand-rolled wells are bhp - val or qWs + qOs + qGs - val),
 * status[nw] (1 open, 0 shut: rates zero, bhp = mean perforation
 * pressure), and rscale[3], the rate-closure rows' per-phase scale (the
 * facility's rhoScale, facility_model.py:2239-2263; 1 for the hand-rolled
 * model). */
int pmx_bo_set_controls(pmx_bo* h, const double* coef, const double* val, const int32_t* status,
                        const double* rscale);
/* Pressure drop from each well's bhp datum to each perforation (host,
 * nperf, connection_pressure_drop, facility_model.py:2294-2389), held
 * until changed (zero for the hand-rolled model, _well :383-436). */
int pmx_bo_set_connection_drops(pmx_bo* h, const double* cdp);
/* Reservoir volume rates of every perforation at the last pmx_bo_assemble
 * (host, 3*nperf: a producing perforation's phase rates, an injecting
 * one's total rate times compi; the facility's cq_vol,
 * facility_model.py:2511-2528), which a converged step hands to the next
 * one's wellbore column (generic_black_oil_model.py:569-601). */
int pmx_bo_perf_volume_rates(pmx_bo* h, double* out);
/* The Newton update rule of pmx_bo_update: 0 the hand-rolled model's
 * (update_state :439-470), 1 update_state_black_oil_generic's
 * (pymrst/modules/ad_blackoil/utils/update_state_black_oil_generic.py:10-104:
 * pressure change capped at dp_max_rel |p|, rs change at drs_max, the
 * saturation changes scaled together to at most ds_max). */
int pmx_bo_set_update_rule(pmx_bo* h, int32_t rule, double dp_max_rel, double drs_max);
/* Phase densities (b_w rhoS_w, b_o (rhoS_o + rs rhoS_g), b_g rhoS_g) and
 * mobilities of every cell at a device state, into out (device, 6*nc per
 * cell in that order): the facility's Density and Mobility at state0
 * (facility_model.py:2223-2226). */
int pmx_bo_cell_densities(pmx_bo* h, const double* p, const double* s, const double* rs, const int32_t* status,
                          double* out);
/* y = J x for the last Jacobian (device). */
int pmx_bo_matvec(pmx_bo* h, const double* x, double* y);
/* Block pattern (host) and 3x3 block values (device) of the last Jacobian. */
int pmx_bo_nnzb(pmx_bo* h, int32_t* nnzb);
int pmx_bo_export_pattern(pmx_bo* h, int32_t* rowptr, int32_t* col);
int pmx_bo_export_blocks(pmx_bo* h, double* vals);
/* Kernel timing as pmx_ow_profile / pmx_ow_profile_read. */
int pmx_bo_profile(pmx_bo* h, int32_t enable);
int pmx_bo_profile_read(pmx_bo* h, double* out4);
/* pmx_ow_launch_shapes for the black-oil model's solver. */
int pmx_bo_launch_shapes(pmx_bo* h, int32_t* out, int32_t cap, int32_t* count);

/* ---- corner-point grids: MRST's processGRDECL topology
 *      (pymrst/gridprocessing/process_grdecl.py:352-580) on the host and
 *      computeGeometry (pymrst/gridprocessing/compute_geometry.py:211-325)
 *      on the device */
typedef struct pmx_cp pmx_cp;

/* Process a corner-point deck (host arrays): coord[6*(nx+1)*(ny+1)] pillar
 * end points (COORD), zcorn[8*nx*ny*nz] corner depths (ZCORN, in MRST's
 * order), actnum[nx*ny*nz] (ACTNUM; null: all active).  Faults (pillar
 * pairs whose cell stacks do not match) are split into the overlaps of the
 * cells on either side; zero-thickness layers join the cells above and
 * below them; inactive and pinched cells are dropped. */
int pmx_cp_process(int32_t nx, int32_t ny, int32_t nz, const double* coord, const double* zcorn,
                   const int32_t* actnum, pmx_cp** out);
int pmx_cp_destroy(pmx_cp* h);
/* out4 = {cells, faces, nodes, face-node entries}. */
int pmx_cp_sizes(pmx_cp* h, int32_t* out4);
/* The grid (host arrays, sized by pmx_cp_sizes): index_map[nc] (Cartesian
 * index i + nx (j + ny k) of each cell), neighbors[2*nf] (-1 outside; a
 * face's area vector points from its first cell to its second), tag[nf]
 * (1 i-face, 2 j-face, 3 k-face; +3 on a fault), node_pos[nf+1] and
 * face_nodes[entries] (faces.nodePos / faces.nodes), coords[3*nn],
 * face_pos[nc+1] and cell_faces[2*internal+boundary] (cells.facePos /
 * cells.faces), cell_dims[3*nc] (extents of each cell's corner points). */
int pmx_cp_topology(pmx_cp* h, int32_t* index_map, int32_t* neighbors, int32_t* tag, int32_t* node_pos,
                    int32_t* face_nodes, double* coords, int32_t* face_pos, int32_t* cell_faces, double* cell_dims);
/* computeGeometry on the device (device arrays): face_normals[3*nf] (area
 * vectors), face_areas[nf], face_centroids[3*nf], cell_volumes[nc],
 * cell_centroids[3*nc]. */
int pmx_cp_geometry(pmx_cp* h, int32_t device, void* hip_stream, double* face_normals, double* face_areas,
                    double* face_centroids, double* cell_volumes, double* cell_centroids);

/* ---- incompressible TPFA pressure solve (incomp_tpfa_ad,
 *      pymrst/modules/incomp/incomp_tpfa.py:252-389) and computeTrans
 *      (pymrst/solvers/compute_trans.py:11-29) */
typedef struct pmx_incomp pmx_incomp;

int pmx_incomp_create(int32_t device, void* hip_stream, pmx_incomp** out);
int pmx_incomp_destroy(pmx_incomp* h);

/* Grid connectivity and geometry (host): face_pos[nc+1] and hf_face[nhf]
 * are MRST's cells.facePos / cells.faces, face_nb[2*nf] its faces.neighbors
 * (-1 outside), cell_centroids[3*nc], face_centroids[3*nf]. */
int pmx_incomp_set_grid(pmx_incomp* h, int32_t nc, int32_t nf, const int32_t* face_pos, const int32_t* hf_face,
                        const int32_t* face_nb, const double* cell_centroids, const double* face_centroids);

/* computeTrans on the device: one-sided transmissibility per half-face into
 * T[nhf] from face_normals[3*nf] and perm[nk*nc] (nk = 1, 3 or 6; all device). */
int pmx_compute_trans(pmx_incomp* h, const double* face_normals, int32_t nk, const double* perm, double* T);

/* Pressure solve.  Device: T[nhf] (one-sided transmissibilities), mob[nc]
 * (total mobility), omega[nc] (mobility-weighted density), and the outputs
 * p[nc], flux[nf], face_pressure[nf].  Host: gravity[3] (NULL = off),
 * boundary conditions (bc_type 0 pressure, 1 flux), sources, wells
 * (perf_ptr[nw+1], perf_cells, WI, wtype 0 bhp / 1 rate, wval, wdp = the
 * connection pressure drops) and the outputs well_p[nw], well_flux[nperf]. */
int pmx_incomp_solve(pmx_incomp* h, const double* T, const double* mob, const double* omega, const double* gravity,
                     int32_t nbc, const int32_t* bc_face, const int32_t* bc_type, const double* bc_value, int32_t nsrc,
                     const int32_t* src_cell, const double* src_rate, int32_t nw, const int32_t* perf_ptr,
                     const int32_t* perf_cells, const double* WI, const int32_t* wtype, const double* wval,
                     const double* wdp, const pmx_solver_options* o, double* p, double* flux, double* face_pressure,
                     double* well_p, double* well_flux, pmx_solve_info* info);

/* The stages of pmx_incomp_solve, for the reference's MatrixOutput and
 * LinSolve options (incomp_tpfa.py:55-63, 172-184):
 * pmx_incomp_assemble builds the system A x = rhs of incomp_tpfa
 * (:120-171, the null-space pin included) from the same inputs as
 * pmx_incomp_solve, plus periodic-boundary pressure jumps (bcp: nbcp faces
 * and values added across those faces, :113-118);
 * pmx_incomp_nnz / pmx_incomp_export copy it to host arrays: the cell
 * matrix in CSR (rowptr[nc+1], col[nnz], val[nnz]), per perforation the
 * cell-row/well-column and well-row/cell-column entries Bp[nperf],
 * Cp[nperf], the well diagonal D[nw] and rhs[nc+nw] (any pointer may be
 * NULL); pmx_incomp_solve_assembled solves it on the device into x
 * (device, nc + nw: cell pressures, then well pressures);
 * pmx_incomp_reconstruct computes pressures, face pressures, fluxes and
 * the wells' pressures and connection fluxes from any solution x (device)
 * of the last assembled system (:186-210). */
int pmx_incomp_assemble(pmx_incomp* h, const double* T, const double* mob, const double* omega, const double* gravity,
                        int32_t nbc, const int32_t* bc_face, const int32_t* bc_type, const double* bc_value,
                        int32_t nsrc, const int32_t* src_cell, const double* src_rate, int32_t nbcp,
                        const int32_t* bcp_face, const double* bcp_value, int32_t nw, const int32_t* perf_ptr,
                        const int32_t* perf_cells, const double* WI, const int32_t* wtype, const double* wval,
                        const double* wdp);
int pmx_incomp_nnz(pmx_incomp* h, int32_t* nnz);
int pmx_incomp_export(pmx_incomp* h, int32_t* rowptr, int32_t* col, double* val, double* Bp, double* Cp, double* D,
                      double* rhs);
int pmx_incomp_solve_assembled(pmx_incomp* h, const pmx_solver_options* o, double* x, pmx_solve_info* info);
int pmx_incomp_reconstruct(pmx_incomp* h, const double* x, double* p, double* flux, double* face_pressure,
                           double* well_p, double* well_flux);

/* ---- a caller's linear system -----------------------------------------
 * Replaces the reference's linear-solver seam: CPRSolver.solve(jac, rhs,
 * model, state, dt) (pymrst/modules/ad_core/solvers/linear_solvers.py:426-545)
 * as called from NonLinearSolver._solve_linear (simulators/nonlinear_solver.py
 * :383-423) and TwoPhaseOilWaterModelAD._solve_increment
 * (two_phase_oil_water_model.py:597-603), for a Jacobian assembled by the
 * caller.  Cell unknowns cell-major in B x B blocks (B = 2 oil-water, 3
 * black-oil), then M = B + 1 unknowns per well (B surface rates, bhp).
 * Host: rowptr[nc+1], col[nnzb] (ascending per row, every diagonal block
 * present), perf_ptr[nw+1], perf_cells[nperf]. */
typedef struct pmx_sys pmx_sys;
int pmx_sys_create(int32_t device, void* hip_stream, int32_t B, int32_t nc, const int32_t* rowptr,
                   const int32_t* col, int32_t nw, const int32_t* perf_ptr, const int32_t* perf_cells,
                   pmx_sys** out);
int pmx_sys_destroy(pmx_sys* h);
/* Device: val[nnzb*B*B] (row-major blocks), Bp[nperf*B*M] (a perforated
 * cell's rows against its well's unknowns), Cp[nperf*M*B] (the well's rows
 * against the cell), Dp[nw*M*M].  The first call also builds the pressure
 * aggregation and the lines from |a_pp| of these values. */
int pmx_sys_set_values(pmx_sys* h, const double* val, const double* Bp, const double* Cp, const double* Dp);
/* y = J x (device vectors of nc*B + nw*M). */
int pmx_sys_matvec(pmx_sys* h, const double* x, double* y);
/* J x = rhs by CPR-FGMRES; w[B] (host) are the quasi-IMPES equation weights
 * of CPRSolver._weighting (linear_solvers.py:155-190).  Returns 0 converged,
 * 1 not converged, <0 error. */
int pmx_sys_solve(pmx_sys* h, const pmx_solver_options* o, const double* w, const double* rhs, double* x,
                  pmx_solve_info* info);

/* The CPR elliptic-solver injection of the reference
 * (linear_solvers.py:124-131, 253-256: elliptic_solver.solve_linear_system
 * (A_p, r)): fn(user, n, b, x) solves the pressure system A_p x = b for
 * the n cells (host arrays; A_p = W J P with the quasi-IMPES weights of
 * pmx_sys_solve) once per preconditioner application, in place of the
 * device AMG V-cycle, and returns 0 (nonzero fails the solve).  fn NULL
 * returns to the AMG.  Applications then run without graph replay. */
typedef int (*pmx_pressure_solver)(void* user, int32_t n, const double* b, double* x);
int pmx_sys_set_pressure_solver(pmx_sys* h, pmx_pressure_solver fn, void* user);

#ifdef __cplusplus
}
#endif
#endif
