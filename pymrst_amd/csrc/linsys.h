// Device linear systems of the TPFA Newton step: a cell-block BCSR matrix
// coupled to a small well tail, the CPR two-stage preconditioner (pressure
// aggregation AMG + cell block-Jacobi) and right-preconditioned flexible
// GMRES.  All vectors are device arrays in the layout
//   [cell 0: B unknowns][cell 1: B unknowns] ... [well 0: M unknowns] ...
// i.e. cell-major interleaved cell unknowns followed by well-major well
// unknowns.
#pragma once
#include "pmx_internal.h"

namespace pmx {

constexpr int kMaxB = 3;
constexpr int kMaxM = 4;
constexpr int kMaxMix = 3;

// Jacobian of one Newton iteration.
//   J = [ A   B ]   A: nc x nc blocks of BxB (BCSR)
//       [ C   D ]   D: block-diagonal, one MxM block per well
// Well coupling is stored per perforation j (cell c_j, well w_j):
//   Bp[j] (BxM)  d cell-eq(c_j) / d wellvars(w_j)
//   Cp[j] (MxB)  d wellrow(w_j) / d cellvars(c_j)
//   Um[j] (BxX), V[j] (XxB): rank-X coupling between perforations of one well
//     (mixture terms): A += sum_w sum_{j,i in w} Um[j] V[i] at block (c_j, c_i)
struct BlockSystem {
  int B = 2, M = 3, X = 2;
  int nc = 0, nnzb = 0;
  int nw = 0, nperf = 0;
  bool has_mix = false;
  DArr<int> rowptr, col, diag;        // BCSR pattern
  DArr<double> val;                   // nnzb * B * B (row-major blocks)
  // wells
  DArr<int> perf_cell, perf_well, well_ptr;  // perforations grouped by well
  DArr<int> pc_cell, pc_ptr, pc_perf;        // unique perforated cells -> perforations
  DArr<int> pc_of;                           // cell -> index into pc_cell, or -1
  int npc = 0;
  DArr<double> Bp, Cp, Um, V, Dp;
  size_t n() const { return (size_t)nc * B + (size_t)nw * M; }
  // lanes per block row of the SpMV: 0 until chosen (8 without autotuning;
  // with it, the fastest of 4, 8, 16 timed at the first product)
  mutable int spmv_lpr = 0;
  mutable int autotune = 0;
  void set_pattern(const HostPattern& pat, hipStream_t s);
  void set_wells(int nw_, const std::vector<int>& perf_ptr, const std::vector<int>& perf_cells, hipStream_t s);
};

// Pairs of HIP events recorded around one kernel launch on its stream; the
// elapsed times are summed when read.
struct EventTimer {
  bool on = false;
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  void start(hipStream_t s);
  void stop(hipStream_t s);
  // the next (start, stop) pair for a launch that stamps them itself
  // (hipExtLaunchKernelGGL: the dispatch's own begin and end times, no
  // separate event packets); false when the timer is off
  bool pair(hipEvent_t* a, hipEvent_t* b);
  double total_ms(int* count);   // synchronises on the last event
  void reset() { used = 0; }
  ~EventTimer();
};

// y = J x  (double); the BCSR kernel launch is timed by t when t->on.
void system_matvec(const BlockSystem& S, const double* x, double* y, double* scratch_h, hipStream_t s,
                   EventTimer* t = nullptr);

struct LinearSolverOptions {
  double rtol = 1e-8;
  int restart = 50;
  int maxiter = 200;
  int precision = 64;        // 64: FP64 preconditioner; 32: FP32 preconditioner
  int use_pressure_stage = 1;
  int pre_smooth = 2, post_smooth = 2;
  double jacobi_omega = 1.0;  // AMG Jacobi damping (1.0 measured best on the oil-water and black-oil cases)
  double coarse_alpha = 1.5; // coarse-correction scaling (over-correction for unsmoothed aggregation)
  int amg_cycles = 1;
  double theta = 0.5;
  int coarse_size = 128;
  int pairwise_passes = 0;
  int reorthogonalize = 1;
  int verbose = 0;
  int local_stage = 1;       // 0: cell block-Jacobi; 1: coloured line Gauss-Seidel
  int local_sweeps = 4;      // Gauss-Seidel sweeps of the line local stage
  int use_graphs = 1;        // replay the preconditioner application as a captured HIP graph
  int coarse_block_rows = 0;     // AMG levels up to this size run in one workgroup (the coarsest always does)
  int line_segments = 0;     // lines of 32+ cells solved by segments with spikes (measured slower: off)
  int autotune = 0;          // time the launch-shape candidates of each AMG level and of the SpMV at the
                             // first setup of each size class, keep the fastest (process-wide table)
  int level_shape = -1;      // >= 0: every level above the coarse cycle takes this candidate shape
                             // (0 ELL, 1-3 CSR with 4/8/16 lanes, 4-6 the same folded), for tests
  int dot_shape = 0;         // > 0: the projections take this shape (2, 4, 8; 104, 108), for tests
};

struct SolveInfo {
  int iterations = 0;
  int restarts = 0;
  double rel_residual = 0.0;
  double setup_ms = 0.0, solve_ms = 0.0;
  int levels = 0;
  int coarse_n = 0;
  int reorthogonalizations = 0;
};

class CprSolver;   // opaque (linsys.hip)

CprSolver* cpr_create();
void cpr_destroy(CprSolver*);
// Static aggregation hierarchy from pressure-graph weights per BCSR entry of S.
void cpr_build_hierarchy(CprSolver*, const BlockSystem& S, const std::vector<double>& weights,
                         const LinearSolverOptions& o, hipStream_t s);
bool cpr_has_hierarchy(const CprSolver*);
// Lines for the local stage (line_ptr[nlines+1], cells); without them lines
// are detected from the strength weights.
void cpr_set_lines(CprSolver*, int nlines, const int* ptr, const int* cells);
// A host solver for the pressure stage, x = A_p^{-1} b (n cells, host
// arrays), called once per preconditioner application in place of the AMG
// V-cycle; fn == nullptr returns to the AMG.  Returns 0 on success.
typedef int (*PressureSolveFn)(void* user, int n, const double* b, double* x);
void cpr_set_pressure_solver(CprSolver*, PressureSolveFn fn, void* user);
// Independent segments (ensemble members): cells [cell_ptr[s], cell_ptr[s+1])
// and wells [well_ptr[s], well_ptr[s+1]) form segment s; FGMRES then keeps
// one Krylov recurrence per segment.
void cpr_set_segments(CprSolver*, int nseg, const int* cell_ptr, const int* well_ptr, int B, int M, int nc, int nw);
// Numeric setup for the current values of S (pressure weights: w_dev[B]).
void cpr_setup(CprSolver*, const BlockSystem& S, const double* w_dev, const LinearSolverOptions& o,
               hipStream_t s);
// Same, with prolongation weights pc_dev (B doubles on the device) spreading
// the pressure correction over a block's unknowns; nullptr means e_0.
void cpr_setup(CprSolver*, const BlockSystem& S, const double* w_dev, const double* pc_dev,
               const LinearSolverOptions& o, hipStream_t s);
// Solve J x = b (FGMRES); returns 0 converged, 1 not converged.
int cpr_solve(CprSolver*, const BlockSystem& S, const double* b, double* x, const LinearSolverOptions& o,
              SolveInfo& info, hipStream_t s);
// z = M^{-1} r (one preconditioner application), exposed for tests.
void cpr_apply(CprSolver*, const BlockSystem& S, const double* r, double* z, const LinearSolverOptions& o,
               hipStream_t s);

void cpr_profile(CprSolver*, bool on);
// out[0..3] = {spmv launches, spmv ms, preconditioner applications, preconditioner ms}
void cpr_profile_read(CprSolver*, double* out);
// line-solve launches timed in eager applications: count, ms, algorithmic bytes
void cpr_profile_line_read(CprSolver*, double* out3);
// The launch shapes in use: out[0] = the SpMV's lanes per block row (0 before
// the first product), out[1] = levels reported, then per level above the
// coarse cycle {rows, ELL copy used, lanes per CSR row, folded, tuned}, then
// the projections' shape (elements per thread; +100 for the LDS-sum form;
// 0 before the first solve);
// returns the number of ints written (at most cap).
int cpr_launch_shapes(CprSolver* C, const BlockSystem& S, int* out, int cap);
// entries of the process-wide autotuning table / empty it
int tune_table_size();
void tune_table_clear();
int tune_table_save(const char* path);
int tune_table_load(const char* path);

// Helpers used by several sources.
int grid_for(size_t threads, int block = 256);
double device_dot(const double* a, const double* b, size_t n, double* partial, hipStream_t s);

}  // namespace pmx
