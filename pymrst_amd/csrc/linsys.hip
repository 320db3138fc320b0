// Device linear algebra for the TPFA Newton step (see linsys.h).
//
// The algorithm is the reference's CPR-GMRES
// (pymrst/modules/ad_core/solvers/linear_solvers.py:110-545): quasi-IMPES
// weights W (1/b_i at the mean pressure, :155-190), the pressure system
// A_p = W J P (:469-473), an elliptic stage on A_p, a cell block-Jacobi local
// stage with an exact solve of the well tail (:315-363), combined as
//   M^{-1} r = P solve_p(W r) + local(r - J P solve_p(W r))       (:510-516)
// inside GMRES.  This implementation uses right-preconditioned flexible
// GMRES (so the preconditioner may run in FP32 while the Arnoldi process
// stays FP64) and an aggregation AMG V-cycle for solve_p; the reference's
// pyamg smoothed aggregation is not available to pin against (parity is
// stated on the solution of J x = b, as the reference's own tests do,
// linear_solvers.py:639-649).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "linsys.h"

namespace pmx {

template <class T>
void DArr<T>::zero(hipStream_t s) {
  if (n) PMX_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
}
template struct DArr<double>;
template struct DArr<float>;
template struct DArr<int>;
template struct DArr<unsigned char>;
template struct DArr<char>;

int grid_for(size_t threads, int block) { return (int)((threads + block - 1) / block); }

// Input and output vectors of one preconditioner application, read by the
// kernels from device memory so that a captured graph of the application
// serves every Krylov column.
struct PrecIO {
  const double* r;
  double* z;
};

// The pointers read from PrecIO are generic as far as the compiler knows, so
// their stores would be FLAT; a pending FLAT store makes every later wait on
// a load a wait for all outstanding memory operations, which serialises the
// prefetch of a walk.  These casts tell the compiler they are global memory.
template <class T>
using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem<T>* global_ptr(T* p) {
  return (gmem<T>*)p;
}

__global__ void k_set_io(PrecIO* io, const double* r, double* z) {
  io->r = r;
  io->z = z;
}

template <class T>
__global__ void k_fill(size_t n, T v, T* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

template <class T>
__global__ void k_copy(size_t n, const T* __restrict__ a, T* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = a[i];
}

static inline void launch_check() { PMX_HIP(hipGetLastError()); }

// ---------------------------------------------------------------------------
// BlockSystem host helpers
// ---------------------------------------------------------------------------
void BlockSystem::set_pattern(const HostPattern& pat, hipStream_t s) {
  nc = pat.n;
  nnzb = pat.rowptr[nc];
  rowptr.upload(pat.rowptr, s);
  col.upload(pat.col, s);
  diag.upload(pat.diag, s);
  val.alloc((size_t)nnzb * B * B);
}

void BlockSystem::set_wells(int nw_, const std::vector<int>& perf_ptr, const std::vector<int>& perf_cells,
                            hipStream_t s) {
  nw = nw_;
  nperf = nw ? perf_ptr[nw] : 0;
  std::vector<int> pw(nperf);
  for (int w = 0; w < nw; ++w)
    for (int j = perf_ptr[w]; j < perf_ptr[w + 1]; ++j) pw[j] = w;
  well_ptr.upload(perf_ptr.data(), nw + 1, s);
  perf_cell.upload(perf_cells.data(), nperf, s);
  perf_well.upload(pw, s);
  // unique perforated cells (ascending) -> perforations in input order
  std::vector<int> order(nperf);
  for (int j = 0; j < nperf; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return perf_cells[a] < perf_cells[b]; });
  std::vector<int> cells, ptr(1, 0), perf;
  for (int t = 0; t < nperf; ++t) {
    int j = order[t];
    if (cells.empty() || cells.back() != perf_cells[j]) {
      if (!cells.empty()) ptr.push_back((int)perf.size());
      cells.push_back(perf_cells[j]);
    }
    perf.push_back(j);
  }
  if (!cells.empty()) ptr.push_back((int)perf.size());
  npc = (int)cells.size();
  PMX_CHECK(nc > 0, "set the pattern before the wells");
  std::vector<int> of(nc, -1);
  for (int u = 0; u < npc; ++u) {
    PMX_CHECK(cells[u] >= 0 && cells[u] < nc, "perforated cell outside the grid");
    of[cells[u]] = u;
  }
  pc_of.upload(of, s);
  pc_cell.upload(cells, s);
  pc_ptr.upload(ptr, s);
  pc_perf.upload(perf, s);
  Bp.alloc((size_t)std::max(nperf, 1) * B * M);
  Cp.alloc((size_t)std::max(nperf, 1) * M * B);
  Um.alloc((size_t)std::max(nperf, 1) * B * X);
  V.alloc((size_t)std::max(nperf, 1) * X * B);
  Dp.alloc((size_t)std::max(nw, 1) * M * M);
  Bp.zero(s); Cp.zero(s); Um.zero(s); V.zero(s); Dp.zero(s);
}

// ---------------------------------------------------------------------------
// Matrix-vector products
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T wave_sum(T v, int width) {
  for (int off = width / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, width);
  return v;
}

// y_cells = A x_cells.  LPR lanes cooperate on one block row (coalesced block
// reads), then reduce with cross-lane shuffles.
template <int B, int LPR>
__global__ __launch_bounds__(256) void k_bcsr_spmv(int nc, const int* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ x, double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  double acc[B];
#pragma unroll
  for (int r = 0; r < B; ++r) acc[r] = 0.0;
  if (row < nc) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) {
      const int c = col[k];
      const double* a = val + (size_t)k * B * B;
      double xc[B];
#pragma unroll
      for (int q = 0; q < B; ++q) xc[q] = x[(size_t)c * B + q];
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int q = 0; q < B; ++q) acc[r] = fma(a[r * B + q], xc[q], acc[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < B; ++r) acc[r] = wave_sum(acc[r], LPR);
  if (row < nc && lane == 0) {
#pragma unroll
    for (int r = 0; r < B; ++r) y[(size_t)row * B + r] = acc[r];
  }
}

// Well rows of y = J x and the per-well mixture projections h_w = sum_j V_j x_cj.
// One 64-wide block per well.
template <int B, int M, int X>
__global__ __launch_bounds__(64) void k_well_rows(int nw, const int* __restrict__ well_ptr,
                                                  const int* __restrict__ perf_cell,
                                                  const double* __restrict__ Cp, const double* __restrict__ V,
                                                  const double* __restrict__ Dp, const double* __restrict__ x,
                                                  size_t ncB, double* __restrict__ y, double* __restrict__ h) {
  const int w = blockIdx.x;
  const int lane = threadIdx.x;
  double ay[M], ah[X];
#pragma unroll
  for (int r = 0; r < M; ++r) ay[r] = 0.0;
#pragma unroll
  for (int a = 0; a < X; ++a) ah[a] = 0.0;
  for (int j = well_ptr[w] + lane; j < well_ptr[w + 1]; j += 64) {
    const int c = perf_cell[j];
    double xc[B];
#pragma unroll
    for (int q = 0; q < B; ++q) xc[q] = x[(size_t)c * B + q];
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
      for (int q = 0; q < B; ++q) ay[r] += Cp[(size_t)j * M * B + r * B + q] * xc[q];
#pragma unroll
    for (int a = 0; a < X; ++a)
#pragma unroll
      for (int q = 0; q < B; ++q) ah[a] += V[(size_t)j * X * B + a * B + q] * xc[q];
  }
#pragma unroll
  for (int r = 0; r < M; ++r) ay[r] = wave_sum(ay[r], 64);
#pragma unroll
  for (int a = 0; a < X; ++a) ah[a] = wave_sum(ah[a], 64);
  if (lane == 0) {
    const double* xw = x + ncB + (size_t)w * M;
#pragma unroll
    for (int r = 0; r < M; ++r) {
      double s = ay[r];
#pragma unroll
      for (int q = 0; q < M; ++q) s += Dp[(size_t)w * M * M + r * M + q] * xw[q];
      y[ncB + (size_t)w * M + r] = s;
    }
#pragma unroll
    for (int a = 0; a < X; ++a) h[w * X + a] = ah[a];
  }
}

// Perforated cell rows: y_c += sum_j (Bp_j x_wj + Um_j h_wj).
template <int B, int M, int X>
__global__ void k_perf_add(int npc, const int* __restrict__ pc_cell, const int* __restrict__ pc_ptr,
                           const int* __restrict__ pc_perf, const int* __restrict__ perf_well,
                           const double* __restrict__ Bp, const double* __restrict__ Um,
                           const double* __restrict__ x, size_t ncB, const double* __restrict__ h, int has_mix,
                           double* __restrict__ y) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npc) return;
  const int c = pc_cell[u];
  double acc[B];
#pragma unroll
  for (int r = 0; r < B; ++r) acc[r] = 0.0;
  for (int t = pc_ptr[u]; t < pc_ptr[u + 1]; ++t) {
    const int j = pc_perf[t];
    const int w = perf_well[j];
    const double* xw = x + ncB + (size_t)w * M;
#pragma unroll
    for (int r = 0; r < B; ++r) {
#pragma unroll
      for (int q = 0; q < M; ++q) acc[r] += Bp[(size_t)j * B * M + r * M + q] * xw[q];
      if (has_mix)
#pragma unroll
        for (int a = 0; a < X; ++a) acc[r] += Um[(size_t)j * B * X + r * X + a] * h[w * X + a];
    }
  }
#pragma unroll
  for (int r = 0; r < B; ++r) y[(size_t)c * B + r] += acc[r];
}

EventTimer::~EventTimer() {
  for (auto e : ev) (void)hipEventDestroy(e);
}
void EventTimer::start(hipStream_t s) {
  if (!on) return;
  while (ev.size() < used + 2) {
    hipEvent_t e;
    PMX_HIP(hipEventCreate(&e));
    ev.push_back(e);
  }
  PMX_HIP(hipEventRecord(ev[used], s));
}
void EventTimer::stop(hipStream_t s) {
  if (!on) return;
  PMX_HIP(hipEventRecord(ev[used + 1], s));
  used += 2;
}
bool EventTimer::pair(hipEvent_t* a, hipEvent_t* b) {
  if (!on) return false;
  while (ev.size() < used + 2) {
    hipEvent_t e;
    PMX_HIP(hipEventCreate(&e));
    ev.push_back(e);
  }
  *a = ev[used];
  *b = ev[used + 1];
  used += 2;
  return true;
}
double EventTimer::total_ms(int* count) {
  double t = 0.0;
  if (used) PMX_HIP(hipEventSynchronize(ev[used - 1]));
  for (size_t i = 0; i + 1 < used; i += 2) {
    float ms = 0.f;
    PMX_HIP(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
    t += ms;
  }
  if (count) *count = (int)(used / 2);
  return t;
}

// ---------------------------------------------------------------------------
// Launch-shape autotuning: candidates timed with HIP events on the caller's
// stream, the fastest kept per size class in a process-wide table (the
// class: the kind of launch, its element size and block size, n and the
// entries per row each in quarter-octave steps), so that a system of a size
// already met reuses the choice without timing.  Never inside a capture.
// ---------------------------------------------------------------------------
struct TuneKey {
  int kind, elem, nb, nz, extra;
  bool operator<(const TuneKey& o) const {
    return std::tie(kind, elem, nb, nz, extra) < std::tie(o.kind, o.elem, o.nb, o.nz, o.extra);
  }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, int> g_tune;

static int octave4(double v) { return v > 0 ? (int)std::lround(4.0 * std::log2(v)) : -1; }

static bool tune_lookup(const TuneKey& k, int* choice) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return false;
  *choice = it->second;
  return true;
}

static void tune_store(const TuneKey& k, int choice) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune[k] = choice;
}

static bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  PMX_HIP(hipStreamIsCapturing(s, &st));
  return st != hipStreamCaptureStatusNone;
}

// Mean milliseconds of `reps` calls of f on s after one untimed call.
template <class F>
static float time_launches(hipStream_t s, int reps, F&& f) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PMX_HIP(hipEventCreate(&e0));
  PMX_HIP(hipEventCreate(&e1));
  f();
  PMX_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) f();
  PMX_HIP(hipEventRecord(e1, s));
  PMX_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  PMX_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

// A candidate replaces the fixed default shape only when it is faster by
// more than this fraction (choices stay put under timing noise).
constexpr float kTuneMargin = 0.03f;

// The fastest of `times` (candidate i at times[i]) unless the default
// candidate `def` is within kTuneMargin of it.
static int tune_pick(const std::vector<float>& times, int def) {
  int best = def;
  for (int i = 0; i < (int)times.size(); ++i)
    if (times[i] < times[best]) best = i;
  return times[best] < (1.0f - kTuneMargin) * times[def] ? best : def;
}

int tune_table_size() {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  return (int)g_tune.size();
}

void tune_table_clear() {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
}

// The table as text, one entry per line: kind, element size, rows class,
// entries-per-row class, extra, choice.
int tune_table_save(const char* path) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  FILE* f = std::fopen(path, "w");
  PMX_CHECK(f != nullptr, std::string("cannot write the tuning table to ") + path);
  for (const auto& [k, c] : g_tune) std::fprintf(f, "%d %d %d %d %d %d\n", k.kind, k.elem, k.nb, k.nz, k.extra, c);
  const bool ok = std::fclose(f) == 0;
  PMX_CHECK(ok, std::string("cannot write the tuning table to ") + path);
  return (int)g_tune.size();
}

// Entries read from a table written by tune_table_save replace those of the
// same class; returns the number read.
int tune_table_load(const char* path) {
  FILE* f = std::fopen(path, "r");
  PMX_CHECK(f != nullptr, std::string("cannot read the tuning table ") + path);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  int n = 0;
  TuneKey k{};
  int c = 0, got = 0;
  while ((got = std::fscanf(f, "%d %d %d %d %d %d", &k.kind, &k.elem, &k.nb, &k.nz, &k.extra, &c)) == 6) {
    g_tune[k] = c;
    ++n;
  }
  const bool clean = got == EOF;
  std::fclose(f);
  PMX_CHECK(clean, std::string("malformed tuning table ") + path);
  return n;
}

template <int B, int LPR>
static void spmv_launch(const BlockSystem& S, const double* x, double* y, hipStream_t s, hipEvent_t e0,
                        hipEvent_t e1) {
  if (e0)
    hipExtLaunchKernelGGL((k_bcsr_spmv<B, LPR>), dim3(grid_for((size_t)S.nc * LPR)), dim3(256), 0, s, e0, e1, 0,
                          S.nc, S.rowptr.p, S.col.p, S.val.p, x, y);
  else
    hipLaunchKernelGGL((k_bcsr_spmv<B, LPR>), dim3(grid_for((size_t)S.nc * LPR)), dim3(256), 0, s, S.nc,
                       S.rowptr.p, S.col.p, S.val.p, x, y);
  launch_check();
}

template <int B>
static void spmv_any(int lpr, const BlockSystem& S, const double* x, double* y, hipStream_t s,
                     hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (lpr == 4) spmv_launch<B, 4>(S, x, y, s, e0, e1);
  else if (lpr == 16) spmv_launch<B, 16>(S, x, y, s, e0, e1);
  else spmv_launch<B, 8>(S, x, y, s, e0, e1);
}

// The SpMV's lanes per block row: 8, or with autotuning the fastest of 4, 8
// and 16 on this system (y is overwritten by the product that follows).
template <int B>
static int spmv_shape(const BlockSystem& S, const double* x, double* y, hipStream_t s) {
  if (S.spmv_lpr) return S.spmv_lpr;
  if (!S.autotune || stream_capturing(s)) return 8;
  const TuneKey k{1, B, octave4(S.nc), octave4((double)S.nnzb / std::max(S.nc, 1)), 0};
  int best = 8;
  if (!tune_lookup(k, &best)) {
    const int cand[3] = {4, 8, 16};
    std::vector<float> t;
    for (int lpr : cand) t.push_back(time_launches(s, 5, [&] { spmv_any<B>(lpr, S, x, y, s); }));
    best = cand[tune_pick(t, 1)];
    tune_store(k, best);
  }
  S.spmv_lpr = best;
  return best;
}

template <int B, int M, int X>
static void matvec_t(const BlockSystem& S, const double* x, double* y, double* h, hipStream_t s, EventTimer* t) {
  const size_t ncB = (size_t)S.nc * B;
  if (S.nc > 0) {
    const int lpr = spmv_shape<B>(S, x, y, s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (t && t->pair(&e0, &e1)) spmv_any<B>(lpr, S, x, y, s, e0, e1);
    else spmv_any<B>(lpr, S, x, y, s);
  }
  if (S.nw > 0) {
    hipLaunchKernelGGL((k_well_rows<B, M, X>), dim3(S.nw), dim3(64), 0, s, S.nw, S.well_ptr.p, S.perf_cell.p,
                       S.Cp.p, S.V.p, S.Dp.p, x, ncB, y, h);
    launch_check();
    if (S.npc > 0) {
      hipLaunchKernelGGL((k_perf_add<B, M, X>), dim3(grid_for(S.npc)), dim3(256), 0, s, S.npc, S.pc_cell.p,
                         S.pc_ptr.p, S.pc_perf.p, S.perf_well.p, S.Bp.p, S.Um.p, x, ncB, h, (int)S.has_mix, y);
      launch_check();
    }
  }
}

#define PMX_DISPATCH(S, CALL)                                                     \
  do {                                                                            \
    if ((S).B == 2 && (S).M == 3 && (S).X == 2) {                                 \
      constexpr int B_ = 2, M_ = 3, X_ = 2;                                       \
      CALL;                                                                       \
    } else if ((S).B == 1 && (S).M == 1 && (S).X == 1) {                          \
      constexpr int B_ = 1, M_ = 1, X_ = 1;                                       \
      CALL;                                                                       \
    } else if ((S).B == 3 && (S).M == 4 && (S).X == 3) {                          \
      constexpr int B_ = 3, M_ = 4, X_ = 3;                                       \
      CALL;                                                                       \
    } else {                                                                      \
      throw std::invalid_argument("unsupported block configuration");             \
    }                                                                             \
  } while (0)

void system_matvec(const BlockSystem& S, const double* x, double* y, double* h, hipStream_t s, EventTimer* t) {
  PMX_DISPATCH(S, (matvec_t<B_, M_, X_>(S, x, y, h, s, t)));
}

// ---------------------------------------------------------------------------
// Vector kernels (FP64, deterministic two-stage reductions)
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 256;   // reductions: threads per block
constexpr int kRedElems = 4;       // elements per thread (a block covers one tile)
constexpr int kRedTile = kRedThreads * kRedElems;
constexpr int kRedGroup = 8;       // dot products formed per pass over a tile

static inline int red_blocks(size_t n) { return (int)((n + kRedTile - 1) / kRedTile); }

// partial[blk * nd + d] = sum over tile blk of V_d . w  (d < nv) and w.w
// (d == nv).  w stays in registers; each group of kRedGroup dots keeps
// kRedGroup * kRedElems loads in flight per thread.
__global__ __launch_bounds__(kRedThreads) void k_multidot(size_t n, int nv, const double* __restrict__ V,
                                                          size_t ld, const double* __restrict__ w,
                                                          double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64][kRedGroup];
  const int nd = nv + 1;
  const size_t base = (size_t)blockIdx.x * kRedTile + threadIdx.x;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double wv[kRedElems];
#pragma unroll
  for (int e = 0; e < kRedElems; ++e) {
    const size_t i = base + (size_t)e * kRedThreads;
    wv[e] = i < n ? w[i] : 0.0;
  }
  for (int d0 = 0; d0 < nd; d0 += kRedGroup) {
    double acc[kRedGroup];
#pragma unroll
    for (int q = 0; q < kRedGroup; ++q) {
      const int d = d0 + q;
      double a = 0.0;
      if (d < nv) {
#pragma unroll
        for (int e = 0; e < kRedElems; ++e) {
          const size_t i = base + (size_t)e * kRedThreads;
          if (i < n) a = fma(V[(size_t)d * ld + i], wv[e], a);
        }
      } else if (d == nv) {
#pragma unroll
        for (int e = 0; e < kRedElems; ++e) a = fma(wv[e], wv[e], a);
      }
      acc[q] = wave_sum(a, 64);
    }
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < kRedGroup; ++q) sh[wid][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < kRedGroup && d0 + (int)threadIdx.x < nd) {
      double t = 0.0;
#pragma unroll
      for (int v = 0; v < kRedThreads / 64; ++v) t += sh[v][threadIdx.x];
      partial[(size_t)blockIdx.x * nd + d0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// out[d] = sum_blk partial[blk*nd + d]: one 256-wide block per d, fixed
// order within each thread's stride and a fixed shuffle tree, so the result
// does not depend on timing.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
  return s;
}

__global__ __launch_bounds__(256) void k_reduce_partials(int nblk, int nd, const double* __restrict__ partial,
                                                         double* __restrict__ out) {
  __shared__ double sh[4];
  const int d = blockIdx.x;
  if (d >= nd) return;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += partial[(size_t)b * nd + d];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[d] = s;
}

// w -= sum_d coef[d] V_d ; partial[blk] = sum of w^2 over tile blk.
__global__ __launch_bounds__(kRedThreads) void k_multiaxpy_norm(size_t n, int nv, const double* __restrict__ V,
                                                                size_t ld, const double* __restrict__ coef,
                                                                double* __restrict__ w, double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64];
  const size_t base = (size_t)blockIdx.x * kRedTile + threadIdx.x;
  double wv[kRedElems];
#pragma unroll
  for (int e = 0; e < kRedElems; ++e) {
    const size_t i = base + (size_t)e * kRedThreads;
    wv[e] = i < n ? w[i] : 0.0;
  }
  for (int d = 0; d < nv; ++d) {
    const double c = coef[d];
#pragma unroll
    for (int e = 0; e < kRedElems; ++e) {
      const size_t i = base + (size_t)e * kRedThreads;
      if (i < n) wv[e] = fma(-c, V[(size_t)d * ld + i], wv[e]);
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < kRedElems; ++e) {
    const size_t i = base + (size_t)e * kRedThreads;
    if (i < n) {
      w[i] = wv[e];
      acc = fma(wv[e], wv[e], acc);
    }
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// w = (w - sum_d coef[d] V_d) * scale
__global__ void k_multiaxpy_scale(size_t n, int nv, const double* __restrict__ V, size_t ld,
                                  const double* __restrict__ coef, double scale, double* __restrict__ w) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double wi = w[i];
  for (int d = 0; d < nv; ++d) wi = fma(-coef[d], V[(size_t)d * ld + i], wi);
  w[i] = wi * scale;
}

// Two right-hand sides against the basis, for delayed reorthogonalisation:
// partial[blk][0..nv)    = V_d . a   (a: the lagged vector in slot nv)
// partial[blk][nv..2nv)  = V_d . w
// partial[blk][2nv..+3)  = a . a, a . w, w . w
// Each basis vector is read once.
// E elements per thread: a block reduces each dot over kRedThreads * E
// elements, so the cross-lane sums and block barriers, one set per group of
// kDot2Group basis vectors, are spread over E loads per thread and vector.
constexpr int kDot2Group = kRedGroup / 2;
constexpr int kDot2Elems = 8;      // elements per thread of the single-system solve (1% faster than 4, 16 slower)
// Tiles are uniform (kRedThreads * E, the last one short) when tstart is
// null, else tile b covers [tstart[b], tstart[b] + tlen[b]) (segments: no
// tile straddles two members; tlen <= kRedThreads * E).
template <int E>
__global__ __launch_bounds__(kRedThreads) void k_multidot2(size_t n, int nv, const double* __restrict__ V,
                                                           size_t ld, const double* __restrict__ a,
                                                           const double* __restrict__ w,
                                                           double* __restrict__ partial,
                                                           const int* __restrict__ tstart,
                                                           const int* __restrict__ tlen) {
  __shared__ double sh[kRedThreads / 64][kRedGroup];
  const int nd = 2 * nv + 3;
  constexpr size_t tile = (size_t)kRedThreads * E;
  const size_t t0 = tstart ? (size_t)tstart[blockIdx.x] : (size_t)blockIdx.x * tile;
  const size_t len = tstart ? (size_t)tlen[blockIdx.x] : (n - t0 < tile ? n - t0 : tile);
  const size_t base = t0 + threadIdx.x;
  const size_t end = t0 + len;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double av[E], wv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const size_t i = base + (size_t)e * kRedThreads;
    av[e] = i < end ? a[i] : 0.0;
    wv[e] = i < end ? w[i] : 0.0;
  }
  for (int d0 = 0; d0 < nv; d0 += kDot2Group) {
    double acc[kRedGroup];
#pragma unroll
    for (int q = 0; q < kDot2Group; ++q) {
      const int d = d0 + q;
      double sa = 0.0, sw = 0.0;
      if (d < nv) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const size_t i = base + (size_t)e * kRedThreads;
          if (i < end) {
            const double v = V[(size_t)d * ld + i];
            sa = fma(v, av[e], sa);
            sw = fma(v, wv[e], sw);
          }
        }
      }
      acc[q] = wave_sum(sa, 64);
      acc[kDot2Group + q] = wave_sum(sw, 64);
    }
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < kRedGroup; ++q) sh[wid][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < kRedGroup) {
      const int q = threadIdx.x % kDot2Group, side = threadIdx.x / kDot2Group;
      if (d0 + q < nv) {
        double t = 0.0;
#pragma unroll
        for (int v = 0; v < kRedThreads / 64; ++v) t += sh[v][threadIdx.x];
        partial[(size_t)blockIdx.x * nd + side * nv + d0 + q] = t;
      }
    }
    __syncthreads();
  }
  double aa = 0.0, aw = 0.0, ww = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    aa = fma(av[e], av[e], aa);
    aw = fma(av[e], wv[e], aw);
    ww = fma(wv[e], wv[e], ww);
  }
  aa = wave_sum(aa, 64);
  aw = wave_sum(aw, 64);
  ww = wave_sum(ww, 64);
  if (lane == 0) {
    sh[wid][0] = aa;
    sh[wid][1] = aw;
    sh[wid][2] = ww;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < kRedThreads / 64; ++v) t += sh[v][threadIdx.x];
    partial[(size_t)blockIdx.x * nd + 2 * nv + threadIdx.x] = t;
  }
}

// The same projections with the cross-lane sums done through LDS: each
// pass over kLdsGroup basis vectors leaves 2 kLdsGroup partial dots per
// thread, stored transposed in LDS; then 16 threads reduce each dot (16
// values each, conflict-free reads, then a 16-lane shuffle tree), one
// barrier pair per pass instead of a 64-lane shuffle tree per dot.
constexpr int kLdsGroup = 8;
template <int E>
__global__ __launch_bounds__(kRedThreads) void k_multidot2_lds(size_t n, int nv, const double* __restrict__ V,
                                                               size_t ld, const double* __restrict__ a,
                                                               const double* __restrict__ w,
                                                               double* __restrict__ partial) {
  static_assert(2 * kLdsGroup * 16 == kRedThreads, "16 threads per dot of a pass");
  __shared__ double sh[2 * kLdsGroup][kRedThreads];
  __shared__ double s3[kRedThreads / 64][3];
  const int nd = 2 * nv + 3;
  constexpr size_t tile = (size_t)kRedThreads * E;
  const size_t t0 = (size_t)blockIdx.x * tile;
  const size_t len = n - t0 < tile ? n - t0 : tile;
  const size_t base = t0 + threadIdx.x;
  const size_t end = t0 + len;
  const int tid = threadIdx.x;
  double av[E], wv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const size_t i = base + (size_t)e * kRedThreads;
    av[e] = i < end ? a[i] : 0.0;
    wv[e] = i < end ? w[i] : 0.0;
  }
  const int q = tid >> 4, r = tid & 15;   // this thread's dot in the reduction, and its slice
  for (int d0 = 0; d0 < nv; d0 += kLdsGroup) {
#pragma unroll
    for (int g = 0; g < kLdsGroup; ++g) {
      const int d = d0 + g;
      double sa = 0.0, sw = 0.0;
      if (d < nv) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const size_t i = base + (size_t)e * kRedThreads;
          if (i < end) {
            const double v = V[(size_t)d * ld + i];
            sa = fma(v, av[e], sa);
            sw = fma(v, wv[e], sw);
          }
        }
      }
      sh[g][tid] = sa;
      sh[kLdsGroup + g][tid] = sw;
    }
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += sh[q][k * 16 + r];
    t = wave_sum(t, 16);
    const int side = q / kLdsGroup, g = q % kLdsGroup;
    if (r == 0 && d0 + g < nv) partial[(size_t)blockIdx.x * nd + side * nv + d0 + g] = t;
    __syncthreads();
  }
  double aa = 0.0, aw = 0.0, ww = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    aa = fma(av[e], av[e], aa);
    aw = fma(av[e], wv[e], aw);
    ww = fma(wv[e], wv[e], ww);
  }
  aa = wave_sum(aa, 64);
  aw = wave_sum(aw, 64);
  ww = wave_sum(ww, 64);
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) {
    s3[wid][0] = aa;
    s3[wid][1] = aw;
    s3[wid][2] = ww;
  }
  __syncthreads();
  if (tid < 3) {
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < kRedThreads / 64; ++v) t += s3[v][tid];
    partial[(size_t)blockIdx.x * nd + 2 * nv + tid] = t;
  }
}

// Delayed reorthogonalisation step, one read of the basis:
//   a <- (a - sum_d s_d V_d) * inv_nu                       (coef[0..nv))
//   w <- (w - sum_d c_d V_d - a_coef * a_old) * scale       (coef[nv..2nv))
__global__ void k_dcgs2_update(size_t n, int nv, const double* __restrict__ V, size_t ld,
                               const double* __restrict__ coef, double inv_nu, double a_coef, double scale,
                               double* __restrict__ a, double* __restrict__ w) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double av = a[i];
  double sa = 0.0, sw = 0.0;
  for (int d = 0; d < nv; ++d) {
    const double v = V[(size_t)d * ld + i];
    sa = fma(coef[d], v, sa);
    sw = fma(coef[nv + d], v, sw);
  }
  a[i] = (av - sa) * inv_nu;
  w[i] = (w[i] - sw - a_coef * av) * scale;
}

// x += sum_d coef[d] Z_d
__global__ void k_multiaxpy_add(size_t n, int nv, const double* __restrict__ Z, size_t ld,
                                const double* __restrict__ coef, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xi = x[i];
  for (int d = 0; d < nv; ++d) xi = fma(coef[d], Z[(size_t)d * ld + i], xi);
  x[i] = xi;
}

__global__ void k_scale_copy(size_t n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

__global__ void k_scale_inplace(size_t n, double a, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

__global__ void k_axpby(size_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

double device_dot(const double* a, const double* b, size_t n, double* partial, hipStream_t s) {
  // partial needs red_blocks(n)*2 + 2 doubles
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(kRedThreads), 0, s, n, 1, a, n, b, partial);
  launch_check();
  hipLaunchKernelGGL(k_reduce_partials, dim3(2), dim3(256), 0, s, nb, 2, partial, partial + (size_t)nb * 2);
  launch_check();
  double h[2];
  PMX_HIP(hipMemcpyAsync(h, partial + (size_t)nb * 2, sizeof(h), hipMemcpyDeviceToHost, s));
  PMX_HIP(hipStreamSynchronize(s));
  return h[0];
}

// ---------------------------------------------------------------------------
// CPR preconditioner pieces
// ---------------------------------------------------------------------------
// Pressure matrix A_p = W J P and the prolonged column J P (per block).
// W combines the rows of a block with weights w; P spreads the pressure
// unknown over the block's columns with weights pc (the pressure column,
// pc = e_0, for the forward system).
template <int B, class T>
__global__ void k_pressure_matrix(int nnzb, const double* __restrict__ val, const double* __restrict__ w,
                                  const double* __restrict__ pc, T* __restrict__ Ap, T* __restrict__ Jp) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnzb) return;
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < B; ++a) {
    double v = 0.0;
#pragma unroll
    for (int b = 0; b < B; ++b) v = fma(val[(size_t)k * B * B + a * B + b], pc[b], v);
    s += w[a] * v;
    Jp[(size_t)k * B + a] = (T)v;
  }
  Ap[k] = (T)s;
}

// Small dense inverse (Gauss-Jordan with partial pivoting) in registers.
template <int N>
__device__ bool small_inverse(double a[N][N], double inv[N][N]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) inv[i][j] = (i == j) ? 1.0 : 0.0;
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) scale = fmax(scale, fabs(a[i][j]));
  if (!(scale > 0.0)) return false;
  for (int k = 0; k < N; ++k) {
    int piv = k;
    double best = fabs(a[k][k]);
    for (int i = k + 1; i < N; ++i)
      if (fabs(a[i][k]) > best) { best = fabs(a[i][k]); piv = i; }
    if (!(best > 1e-14 * scale)) {
      // regularise a singular tail block (cf. linear_solvers.py:336-353)
      a[k][k] += 1.4901161193847656e-08 * scale;
      best = fabs(a[k][k]);
      piv = k;
      if (!(best > 0.0)) return false;
    }
    if (piv != k)
      for (int j = 0; j < N; ++j) {
        double t = a[k][j]; a[k][j] = a[piv][j]; a[piv][j] = t;
        t = inv[k][j]; inv[k][j] = inv[piv][j]; inv[piv][j] = t;
      }
    const double d = 1.0 / a[k][k];
    for (int j = 0; j < N; ++j) { a[k][j] *= d; inv[k][j] *= d; }
    for (int i = 0; i < N; ++i) {
      if (i == k) continue;
      const double f = a[i][k];
      if (f == 0.0) continue;
      for (int j = 0; j < N; ++j) { a[i][j] -= f * a[k][j]; inv[i][j] -= f * inv[k][j]; }
    }
  }
  return true;
}

template <int B, class T>
__global__ void k_block_inverse(int nc, const int* __restrict__ diag, const double* __restrict__ val,
                                T* __restrict__ inv) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double a[B][B], r[B][B];
  const double* v = val + (size_t)diag[c] * B * B;
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) a[i][j] = v[i * B + j];
  small_inverse<B>(a, r);
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) inv[(size_t)c * B * B + i * B + j] = (T)r[i][j];
}

template <int M>
__global__ void k_tail_inverse(int nw, const double* __restrict__ Dp, double* __restrict__ inv) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw) return;
  double a[M][M], r[M][M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) a[i][j] = Dp[(size_t)w * M * M + i * M + j];
  small_inverse<M>(a, r);
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) inv[(size_t)w * M * M + i * M + j] = r[i][j];
}

// r_p = W r (cell rows)
template <int B, class T>
__global__ void k_weight_restrict(int nc, const double* __restrict__ w, const PrecIO* __restrict__ io,
                                  T* __restrict__ rp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const gmem<double>* __restrict__ r = global_ptr(io->r);
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < B; ++a) s += w[a] * r[(size_t)c * B + a];
  rp[c] = (T)s;
}

// Stage 2 on the cell rows: z = P xp + inv(D_c) (r - J P xp).
template <int B, class T, int LPR>
__global__ __launch_bounds__(256) void k_cpr_stage2(int nc, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col, const T* __restrict__ Jp,
                                                    const T* __restrict__ xp, const PrecIO* __restrict__ io,
                                                    const T* __restrict__ inv, const double* __restrict__ pc) {
  const gmem<double>* __restrict__ r = global_ptr(io->r);
  gmem<double>* __restrict__ z = global_ptr(io->z);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  double acc[B];
#pragma unroll
  for (int a = 0; a < B; ++a) acc[a] = 0.0;
  if (row < nc) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) {
      const double xc = (double)xp[col[k]];
#pragma unroll
      for (int a = 0; a < B; ++a) acc[a] = fma((double)Jp[(size_t)k * B + a], xc, acc[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < B; ++a) acc[a] = wave_sum(acc[a], LPR);
  if (row < nc && lane == 0) {
    double r2[B];
#pragma unroll
    for (int a = 0; a < B; ++a) r2[a] = r[(size_t)row * B + a] - acc[a];
#pragma unroll
    for (int a = 0; a < B; ++a) {
      double s = pc[a] * (double)xp[row];
#pragma unroll
      for (int b = 0; b < B; ++b) s = fma((double)inv[(size_t)row * B * B + a * B + b], r2[b], s);
      z[(size_t)row * B + a] = s;
    }
  }
}

// Stage 2 on the well rows (one 64-wide block per well); also hp = V P xp.
template <int B, int M, int X, class T>
__global__ __launch_bounds__(64) void k_cpr_wells(int nw, const int* __restrict__ well_ptr,
                                                  const int* __restrict__ perf_cell, const double* __restrict__ Cp,
                                                  const double* __restrict__ V, const double* __restrict__ Dinv,
                                                  const T* __restrict__ xp, const PrecIO* __restrict__ io,
                                                  size_t ncB, double* __restrict__ hp,
                                                  const double* __restrict__ pc) {
  const gmem<double>* __restrict__ r = global_ptr(io->r);
  gmem<double>* __restrict__ z = global_ptr(io->z);
  const int w = blockIdx.x, lane = threadIdx.x;
  double ay[M], ah[X];
#pragma unroll
  for (int q = 0; q < M; ++q) ay[q] = 0.0;
#pragma unroll
  for (int a = 0; a < X; ++a) ah[a] = 0.0;
  for (int j = well_ptr[w] + lane; j < well_ptr[w + 1]; j += 64) {
    const double xc = (double)xp[perf_cell[j]];
#pragma unroll
    for (int q = 0; q < M; ++q)
#pragma unroll
      for (int b = 0; b < B; ++b) ay[q] += Cp[(size_t)j * M * B + q * B + b] * pc[b] * xc;
#pragma unroll
    for (int a = 0; a < X; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) ah[a] += V[(size_t)j * X * B + a * B + b] * pc[b] * xc;
  }
#pragma unroll
  for (int q = 0; q < M; ++q) ay[q] = wave_sum(ay[q], 64);
#pragma unroll
  for (int a = 0; a < X; ++a) ah[a] = wave_sum(ah[a], 64);
  if (lane == 0) {
    double r2[M];
#pragma unroll
    for (int q = 0; q < M; ++q) r2[q] = r[ncB + (size_t)w * M + q] - ay[q];
#pragma unroll
    for (int q = 0; q < M; ++q) {
      double s = 0.0;
#pragma unroll
      for (int p = 0; p < M; ++p) s += Dinv[(size_t)w * M * M + q * M + p] * r2[p];
      z[ncB + (size_t)w * M + q] = s;
    }
#pragma unroll
    for (int a = 0; a < X; ++a) hp[w * X + a] = ah[a];
  }
}

// Mixture-coupling correction of stage 2 on perforated cells:
// z_c -= inv(D_c) sum_j Um_j hp_wj.
template <int B, int X, class T>
__global__ void k_cpr_mixfix(int npc, const int* __restrict__ pc_cell, const int* __restrict__ pc_ptr,
                             const int* __restrict__ pc_perf, const int* __restrict__ perf_well,
                             const double* __restrict__ Um, const double* __restrict__ hp,
                             const T* __restrict__ inv, const PrecIO* __restrict__ io) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= npc) return;
  gmem<double>* __restrict__ z = global_ptr(io->z);
  const int c = pc_cell[u];
  double acc[B];
#pragma unroll
  for (int a = 0; a < B; ++a) acc[a] = 0.0;
  for (int t = pc_ptr[u]; t < pc_ptr[u + 1]; ++t) {
    const int j = pc_perf[t], w = perf_well[j];
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int x = 0; x < X; ++x) acc[a] += Um[(size_t)j * B * X + a * X + x] * hp[w * X + x];
  }
#pragma unroll
  for (int a = 0; a < B; ++a) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < B; ++b) s += (double)inv[(size_t)c * B * B + a * B + b] * acc[b];
    z[(size_t)c * B + a] -= s;
  }
}

// ---------------------------------------------------------------------------
// Line local stage: exact block-tridiagonal solves along lines of strongly
// coupled cells (one thread per line; position-major tables keep the lanes
// of a wave on neighbouring cells).
// ---------------------------------------------------------------------------
template <int B>
__device__ __forceinline__ void blk_load(const double* p, double a[B][B]) {
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) a[i][j] = p[i * B + j];
}

template <int B, class T>
__device__ __forceinline__ void blk_store(T* p, const double a[B][B]) {
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) p[i * B + j] = (T)a[i][j];
}

// Line Gauss-Seidel local stage.  The local system is A y = r1 on the cell
// rows, r1 = r - J P xp - sum Um_j hp_w (the pressure correction and the
// wells' mixture coupling of P xp removed).  Lines of one colour share no
// matrix entry, so a colour's lines are solved together from
//   A_line y_line = r1 - sum_{j off the line} A_cj y_j
// with y_j the newest values: earlier colours of this sweep, later colours
// of the previous sweep (zero in the first sweep).  The colours are swept in
// order, `sweeps` times.
//
// r1 on every cell row (LPR lanes per row), scattered to the row's line
// position; xpl receives xp in the same order.
template <int B, int X, class T, int LPR>
__global__ __launch_bounds__(256) void k_line_r1(int nc, const int* __restrict__ rowptr, const int* __restrict__ col,
                                                 const T* __restrict__ Jp, const T* __restrict__ xp,
                                                 const PrecIO* __restrict__ io, const int* __restrict__ pc_of,
                                                 const int* __restrict__ pc_ptr, const int* __restrict__ pc_perf,
                                                 const int* __restrict__ perf_well, const double* __restrict__ Um,
                                                 const double* __restrict__ hp, const int* __restrict__ pos,
                                                 T* __restrict__ r1l, T* __restrict__ xpl) {
  const gmem<double>* __restrict__ r = global_ptr(io->r);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  double acc[B];
#pragma unroll
  for (int a = 0; a < B; ++a) acc[a] = 0.0;
  if (row < nc) {
    const int e = rowptr[row + 1];
    for (int q = rowptr[row] + lane; q < e; q += LPR) {
      const double xj = (double)xp[col[q]];
#pragma unroll
      for (int a = 0; a < B; ++a) acc[a] = fma((double)Jp[(size_t)q * B + a], xj, acc[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < B; ++a) acc[a] = wave_sum(acc[a], LPR);
  if (row < nc && lane == 0) {
    double r1[B];
#pragma unroll
    for (int a = 0; a < B; ++a) r1[a] = r[(size_t)row * B + a] - acc[a];
    if (pc_of != nullptr) {
      const int u = pc_of[row];
      if (u >= 0)
        for (int m = pc_ptr[u]; m < pc_ptr[u + 1]; ++m) {
          const int j = pc_perf[m], w = perf_well[j];
#pragma unroll
          for (int a = 0; a < B; ++a)
#pragma unroll
            for (int x = 0; x < X; ++x) r1[a] -= Um[(size_t)j * B * X + a * X + x] * hp[w * X + x];
        }
    }
    const size_t at = pos != nullptr ? (size_t)pos[row] : (size_t)row;
#pragma unroll
    for (int a = 0; a < B; ++a) r1l[at * B + a] = (T)r1[a];
    xpl[at] = xp[row];
  }
}

// k_line_r1 on the finest level's sliced-ELL layout (the pressure system
// shares the Jacobian's block pattern, so the level's slot map places the
// pressure column Jp too): one thread per row, every load of a wave
// coalesced; padding slots hold zero and the row's own column.
template <class T, int B>
__global__ void k_to_sell_vec(int nnz, const int* __restrict__ map, const T* __restrict__ v, T* __restrict__ sv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
#pragma unroll
  for (int a = 0; a < B; ++a) sv[(size_t)map[k] * B + a] = v[(size_t)k * B + a];
}

template <int B, int X, class T>
__global__ __launch_bounds__(256) void k_line_r1_sell(int nc, const int* __restrict__ sl_ptr,
                                                      const int* __restrict__ sl_len, const int* __restrict__ sl_col,
                                                      const T* __restrict__ Jps, const T* __restrict__ xp,
                                                      const PrecIO* __restrict__ io, const int* __restrict__ pc_of,
                                                      const int* __restrict__ pc_ptr,
                                                      const int* __restrict__ pc_perf,
                                                      const int* __restrict__ perf_well,
                                                      const double* __restrict__ Um, const double* __restrict__ hp,
                                                      const int* __restrict__ pos, T* __restrict__ r1l,
                                                      T* __restrict__ xpl) {
  const gmem<double>* __restrict__ r = global_ptr(io->r);
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nc) return;
  const int sl = row >> 6;
  const int base = sl_ptr[sl] + (row & 63), len = sl_len[sl];
  double acc[B];
#pragma unroll
  for (int a = 0; a < B; ++a) acc[a] = 0.0;
#pragma unroll 4
  for (int k = 0; k < len; ++k) {
    const size_t q = (size_t)base + 64 * (size_t)k;
    const double xj = (double)xp[sl_col[q]];
#pragma unroll
    for (int a = 0; a < B; ++a) acc[a] = fma((double)Jps[q * B + a], xj, acc[a]);
  }
  double r1[B];
#pragma unroll
  for (int a = 0; a < B; ++a) r1[a] = r[(size_t)row * B + a] - acc[a];
  if (pc_of != nullptr) {
    const int u = pc_of[row];
    if (u >= 0)
      for (int m = pc_ptr[u]; m < pc_ptr[u + 1]; ++m) {
        const int j = pc_perf[m], w = perf_well[j];
#pragma unroll
        for (int a = 0; a < B; ++a)
#pragma unroll
          for (int x = 0; x < X; ++x) r1[a] -= Um[(size_t)j * B * X + a * X + x] * hp[w * X + x];
      }
  }
  const size_t at = pos != nullptr ? (size_t)pos[row] : (size_t)row;
#pragma unroll
  for (int a = 0; a < B; ++a) r1l[at * B + a] = (T)r1[a];
  xpl[at] = xp[row];
}

// Right-hand sides of the lines of colour `color` (one thread per line
// position): r2 = r1 - sum_h Aoff_h y_nbr_h over the couplings off the line;
// in the first sweep only neighbours of earlier colours have a value yet.
// The line results y are kept by table position (ypos), so off_nbr holds the
// neighbours' table positions.
template <int B, class T>
__global__ __launch_bounds__(256) void k_line_rhs(int nl, int lc0, int nlc, int Lc, int color, int first, int noff,
                                                  size_t npos, const int* __restrict__ len,
                                                  const int* __restrict__ off_nbr,
                                                  const unsigned char* __restrict__ off_col,
                                                  const T* __restrict__ aoff, const T* __restrict__ ypos,
                                                  const T* __restrict__ r1l, T* __restrict__ r2l) {
  constexpr int BB = B * B;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int li = (int)(t % nlc), k = (int)(t / nlc);
  const int l = lc0 + li;
  if (k >= Lc || k >= len[l]) return;
  const size_t at = (size_t)k * nl + l;
  double r2[B];
#pragma unroll
  for (int a = 0; a < B; ++a) r2[a] = (double)r1l[at * B + a];
  for (int h = 0; h < noff; ++h) {
    const size_t e = (size_t)h * npos + at;
    const int j = off_nbr[e];
    if (j < 0 || (first && off_col[e] >= color)) continue;
    double yj[B];
#pragma unroll
    for (int b = 0; b < B; ++b) yj[b] = (double)ypos[(size_t)j * B + b];
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) r2[a] -= (double)aoff[e * BB + a * B + b] * yj[b];
  }
#pragma unroll
  for (int a = 0; a < B; ++a) r2l[at * B + a] = (T)r2[a];
}

// Off-line coupling blocks in table order (one thread per plane entry).
template <int B, class T>
__global__ void k_gather_off(size_t ne, const int* __restrict__ off_q, const double* __restrict__ val,
                             T* __restrict__ aoff) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int q = off_q[e];
#pragma unroll
  for (int m = 0; m < B * B; ++m) aoff[e * B * B + m] = q >= 0 ? (T)val[(size_t)q * B * B + m] : (T)0;
}

// Prefetch depth of the line walks below: the next D positions of a walk
// are in flight while its recurrence runs (every load address is known
// without a data-dependent index), and the scheduling barriers keep the
// compiler from sinking those loads next to their use.  One lane per whole
// line (a single LU walk down and back) took 29.9 us per colour at 60x220x85
// against 22.4 us for the twisted form.  With one wave per line group the
// walks are bound by memory latency per step, not by their arithmetic (the
// same recurrence fed from registers runs in under 1 us) nor by the
// prefetch depth (8 and 16 run alike; 8 by a few percent): what shortens a
// step is fewer memory operations in it.
template <int B, class T>
constexpr int line_depth() {
  return B * B * (int)sizeof(T) <= 36 ? 8 : 4;
}

// Twisted ("burn at both ends") block factorization of each line: the top
// half [0, m) is factored downwards (D'_k = D_k - L_k D'_{k-1}^{-1} U_{k-1}),
// the bottom half (m, L) upwards (D''_k = D_k - U_k D''_{k+1}^{-1} L_{k+1}),
// m = L / 2, and the middle pivot takes both couplings.  At every position
// `Lo` holds the coupling towards the previous cell of that half's walk and
// `Up` the coupling towards the next one back (L_k, U_k on top; U_k, L_k
// below), so both halves are solved by the same walk: forward from the line's
// end to the middle, then back out.
template <int B, class T>
__global__ void k_line_factor_tw(int nl, const int* __restrict__ len, const int* __restrict__ nseg,
                                 const int* __restrict__ cell, const int* __restrict__ lo,
                                 const int* __restrict__ up, const int* __restrict__ diag,
                                 const double* __restrict__ val, T* __restrict__ inv, T* __restrict__ Lo,
                                 T* __restrict__ Up) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  const int L = len[l];
  if (L == 0 || (nseg != nullptr && nseg[l] > 0)) return;  // partitioned lines: k_line_factor_seg
  const int m = L / 2;
  double zero[B][B];
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) zero[i][j] = 0.0;
  // D - A P C for the walk step (A: coupling to the previous cell, C: that
  // cell's coupling back)
  auto schur = [&](double (&D)[B][B], const double (&A)[B][B], const double (&P)[B][B], const double (&Cb)[B][B]) {
    double t[B][B];
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < B; ++q) v += A[i][q] * P[q][j];
        t[i][j] = v;
      }
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < B; ++q) v += t[i][q] * Cb[q][j];
        D[i][j] -= v;
      }
  };
  double Ptop[B][B], Utop[B][B], Pbot[B][B], Lbot[B][B];
  for (int k = 0; k < m; ++k) {
    const size_t at = (size_t)k * nl + l;
    double D[B][B];
    blk_load<B>(val + (size_t)diag[cell[at]] * B * B, D);
    if (k > 0) {
      double A[B][B];
      blk_load<B>(val + (size_t)lo[at] * B * B, A);
      schur(D, A, Ptop, Utop);
      blk_store<B, T>(Lo + at * B * B, A);
    } else {
      blk_store<B, T>(Lo + at * B * B, zero);
    }
    small_inverse<B>(D, Ptop);
    blk_store<B, T>(inv + at * B * B, Ptop);
    blk_load<B>(val + (size_t)up[at] * B * B, Utop);      // k < m <= L - 1: a next cell exists
    blk_store<B, T>(Up + at * B * B, Utop);
  }
  for (int k = L - 1; k > m; --k) {
    const size_t at = (size_t)k * nl + l;
    double D[B][B];
    blk_load<B>(val + (size_t)diag[cell[at]] * B * B, D);
    if (k < L - 1) {
      double A[B][B];
      blk_load<B>(val + (size_t)up[at] * B * B, A);
      schur(D, A, Pbot, Lbot);
      blk_store<B, T>(Lo + at * B * B, A);
    } else {
      blk_store<B, T>(Lo + at * B * B, zero);
    }
    small_inverse<B>(D, Pbot);
    blk_store<B, T>(inv + at * B * B, Pbot);
    blk_load<B>(val + (size_t)lo[at] * B * B, Lbot);      // k > m >= 0: a previous cell exists
    blk_store<B, T>(Up + at * B * B, Lbot);
  }
  const size_t at = (size_t)m * nl + l;
  double D[B][B], Di[B][B];
  blk_load<B>(val + (size_t)diag[cell[at]] * B * B, D);
  if (m > 0) {
    double A[B][B];
    blk_load<B>(val + (size_t)lo[at] * B * B, A);
    schur(D, A, Ptop, Utop);
    blk_store<B, T>(Lo + at * B * B, A);
  } else {
    blk_store<B, T>(Lo + at * B * B, zero);
  }
  if (m < L - 1) {
    double A[B][B];
    blk_load<B>(val + (size_t)up[at] * B * B, A);
    schur(D, A, Pbot, Lbot);
    blk_store<B, T>(Up + at * B * B, A);
  } else {
    blk_store<B, T>(Up + at * B * B, zero);
  }
  small_inverse<B>(D, Di);
  blk_store<B, T>(inv + at * B * B, Di);
}

// One colour's line solves with the twisted factors: a wave holds 32 lines;
// lanes 0-31 walk their lines' top halves downwards, lanes 32-63 the bottom
// halves upwards (the same code, positions mirrored), the two lanes of a
// line meet at its middle cell through a lane exchange, and both walk back
// out.  Every load is at a position inside its line (steps past a lane's
// walk are clamped and their results dropped), so lines of a wave may differ
// in length.  The results go to ypos by table position (consecutive lanes,
// consecutive addresses), where the next colour's right-hand sides read
// them; only the last sweep (Z) also writes the output z = x + pc xp by cell,
// since each sweep overwrites every line cell's z and nothing reads z between
// sweeps.
template <int B, class T, int D, bool Z>
__global__ __launch_bounds__(64) void k_line_gs_tw(int nl, int lc0, const int* __restrict__ len,
                                                   const int* __restrict__ cell, const T* __restrict__ inv,
                                                   const T* __restrict__ Lo, const T* __restrict__ Up,
                                                   T* __restrict__ r2l, const T* __restrict__ xpl,
                                                   const PrecIO* __restrict__ io, T* __restrict__ ypos,
                                                   const double* __restrict__ pcw) {
  constexpr int BB = B * B;
  gmem<double>* __restrict__ z = global_ptr(io->z);
  double pc[B];
#pragma unroll
  for (int a = 0; a < B; ++a) pc[a] = pcw[a];
  const int h = threadIdx.x >> 5;                       // 0: top half, 1: bottom half
  const int l = lc0 + blockIdx.x * 32 + (threadIdx.x & 31);
  const int Lw = len[lc0 + blockIdx.x * 32];            // the wave's longest line (lines sorted by length)
  if (Lw == 0) return;                                  // a wave of padding lines only (wave-uniform)
  const int Ll = len[l];
  const int m = Ll / 2;
  const int J = Ll == 0 ? 0 : (h == 0 ? m : Ll - 1 - m);  // walk steps of this lane
  const int Jmax = (Lw / 2 + D - 1) / D * D;            // the longest walk of the wave, padded
  auto pos = [&](int j) -> int {
    const int jj = min(max(j, 0), max(J - 1, 0));
    return Ll == 0 ? 0 : (h == 0 ? jj : Ll - 1 - jj);
  };
  T ri[D][BB], rl[D][BB], rr[D][B];
  auto fload = [&](int j, T (&vi)[BB], T (&vl)[BB], T (&vr)[B]) {
    const size_t at = (size_t)pos(j) * nl + l;
#pragma unroll
    for (int q = 0; q < BB; ++q) {
      vi[q] = inv[at * BB + q];
      vl[q] = Lo[at * BB + q];
    }
#pragma unroll
    for (int a = 0; a < B; ++a) vr[a] = r2l[at * B + a];
  };
#pragma unroll
  for (int u = 0; u < D; ++u) fload(u, ri[u], rl[u], rr[u]);
  double y[B];
#pragma unroll
  for (int a = 0; a < B; ++a) y[a] = 0.0;
  for (int j0 = 0; j0 < Jmax; j0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 + u;
      double rhs[B], yn[B];
#pragma unroll
      for (int a = 0; a < B; ++a) {
        double v = (double)rr[u][a];
#pragma unroll
        for (int q = 0; q < B; ++q) v -= (double)rl[u][a * B + q] * y[q];
        rhs[a] = v;
      }
#pragma unroll
      for (int a = 0; a < B; ++a) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < B; ++q) v += (double)ri[u][a * B + q] * rhs[q];
        yn[a] = v;
      }
      if (j < J) {
        const size_t at = (size_t)pos(j) * nl + l;
#pragma unroll
        for (int a = 0; a < B; ++a) {
          y[a] = yn[a];
          r2l[at * B + a] = (T)yn[a];
        }
      }
      fload(j + D, ri[u], rl[u], rr[u]);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // the middle cell: its pivot takes the top walk's last y through Lo and the
  // bottom walk's through Up
  double yo[B];
#pragma unroll
  for (int a = 0; a < B; ++a) yo[a] = __shfl_xor(y[a], 32);
  double xm[B];
#pragma unroll
  for (int a = 0; a < B; ++a) xm[a] = 0.0;
  if (Ll > 0) {
    const size_t at = (size_t)m * nl + l;
    double rhs[B];
#pragma unroll
    for (int a = 0; a < B; ++a) {
      double v = (double)r2l[at * B + a];
#pragma unroll
      for (int q = 0; q < B; ++q) {
        const double yt = h == 0 ? y[q] : yo[q], yb = h == 0 ? yo[q] : y[q];
        v -= (double)Lo[at * BB + a * B + q] * yt + (double)Up[at * BB + a * B + q] * yb;
      }
      rhs[a] = v;
    }
#pragma unroll
    for (int a = 0; a < B; ++a) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < B; ++q) v += (double)inv[at * BB + a * B + q] * rhs[q];
      xm[a] = v;
    }
    if (h == 0) {
      if (ypos != nullptr)
#pragma unroll
        for (int a = 0; a < B; ++a) ypos[at * B + a] = (T)xm[a];
      if (Z) {
        const int c = cell[at];
        const double xp = (double)xpl[at];
        if (c >= 0)
#pragma unroll
          for (int a = 0; a < B; ++a) z[(size_t)c * B + a] = xm[a] + pc[a] * xp;
      }
    }
  }
  // back out: the forward results y come from r2l, the pressure correction
  // and the cell index only in the last sweep
  T bi[D][BB], bu[D][BB], by[D][B], bx[D];
  int bc[D];
  auto bload = [&](int j, int u) {
    const size_t at = (size_t)pos(j) * nl + l;
#pragma unroll
    for (int q = 0; q < BB; ++q) {
      bi[u][q] = inv[at * BB + q];
      bu[u][q] = Up[at * BB + q];
    }
#pragma unroll
    for (int a = 0; a < B; ++a) by[u][a] = r2l[at * B + a];
    if (Z) {
      bx[u] = xpl[at];
      bc[u] = cell[at];
    }
  };
#pragma unroll
  for (int u = 0; u < D; ++u) bload(Jmax - 1 - u, u);
  double xn[B];
#pragma unroll
  for (int a = 0; a < B; ++a) xn[a] = xm[a];
  for (int j0 = Jmax - 1; j0 >= 0; j0 -= D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 - u;
      double tq[B], xv[B];
#pragma unroll
      for (int a = 0; a < B; ++a) {
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < B; ++q) v += (double)bu[u][a * B + q] * xn[q];
        tq[a] = v;
      }
#pragma unroll
      for (int a = 0; a < B; ++a) {
        double v = (double)by[u][a];
#pragma unroll
        for (int q = 0; q < B; ++q) v -= (double)bi[u][a * B + q] * tq[q];
        xv[a] = v;
      }
      if (j < J) {
#pragma unroll
        for (int a = 0; a < B; ++a) xn[a] = xv[a];
        if (ypos != nullptr) {
          const size_t at = (size_t)pos(j) * nl + l;
#pragma unroll
          for (int a = 0; a < B; ++a) ypos[at * B + a] = (T)xv[a];
        }
        if (Z && bc[u] >= 0) {
          const size_t c = (size_t)bc[u];
#pragma unroll
          for (int a = 0; a < B; ++a) z[c * B + a] = xv[a] + pc[a] * (double)bx[u];
        }
      }
      bload(j - D, u);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// ---------------------------------------------------------------------------
// Partitioned line solves.  A line of L >= 32 cells is cut at ns - 1
// separator cells into ns segments of 15-31 cells (ns = 2, 4 or 8); each
// segment is factored and walked like a line of its own (the twisted form,
// two lanes per segment), which solves A_ii g_i = r_i with the separators
// held at zero.  The spikes W_L = A_ii^{-1} E_first A(first, s_left) and
// W_R = A_ii^{-1} E_last A(last, s_right) give every segment cell's response
// to its separators, so the separators satisfy the block-tridiagonal Schur
// system
//   (A_ss - L_s W_R,k[last] - U_s W_L,k+1[first]) x_s - L_s W_L,k[last] x_s-
//     - U_s W_R,k+1[first] x_s+ = r_s - L_s g_k[last] - U_s g_k+1[first],
// (L_s, U_s the separator's couplings to the cells before and after it),
// factored once per setup by block Thomas, and each segment cell is then
// x = g - W_L x_s_left - W_R x_s_right.  Exact in exact arithmetic, with
// walks of a quarter (ns = 4) of the twisted walk's length and ns times as
// many lanes per line.
// ---------------------------------------------------------------------------
__host__ __device__ inline int line_nseg(int L) {
  const int n = L / 16;
  return n >= 8 ? 8 : (n >= 4 ? 4 : (n >= 2 ? 2 : 1));
}
// separator k (k < ns - 1) of a line of L cells cut into ns segments
__host__ __device__ inline int line_sep(int L, int ns, int k) { return (k + 1) * (L + 1) / ns - 1; }
// first and last cell of segment k
__host__ __device__ inline void line_segment(int L, int ns, int k, int& a, int& b) {
  a = k == 0 ? 0 : line_sep(L, ns, k - 1) + 1;
  b = k == ns - 1 ? L - 1 : line_sep(L, ns, k) - 1;
}

template <int B>
struct Blk {
  double v[B][B];
  __device__ void zero() {
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) v[i][j] = 0.0;
  }
  template <class T>
  __device__ void load(const T* p) {
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) v[i][j] = (double)p[i * B + j];
  }
  template <class T>
  __device__ void store(T* p) const {
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) p[i * B + j] = (T)v[i][j];
  }
};
template <int B>
__device__ __forceinline__ Blk<B> bmul(const Blk<B>& x, const Blk<B>& y) {
  Blk<B> o;
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) s += x.v[i][k] * y.v[k][j];
      o.v[i][j] = s;
    }
  return o;
}
template <int B>
__device__ __forceinline__ Blk<B> bsub(const Blk<B>& x, const Blk<B>& y) {
  Blk<B> o;
#pragma unroll
  for (int i = 0; i < B; ++i)
#pragma unroll
    for (int j = 0; j < B; ++j) o.v[i][j] = x.v[i][j] - y.v[i][j];
  return o;
}
template <int B>
__device__ __forceinline__ Blk<B> binv(const Blk<B>& x) {
  Blk<B> a = x, o;
  small_inverse<B>(a.v, o.v);
  return o;
}

// Twisted factors of every segment of the lines with nseg[l] > 0 (one thread
// per segment), with the coupling back towards the middle premultiplied by
// the pivot's inverse off the middle (P = D'^{-1} U: the backward step is
// g_k = y_k - P_k g_k+1), then the segment's spikes, solved with the stored
// (rounded) factors as the walks will apply them; a segment at a line's end
// has no spike on that side (zero).
template <int B, class T>
__global__ void k_line_factor_seg(int nl, const int* __restrict__ len, const int* __restrict__ nseg,
                                  const int* __restrict__ cell, const int* __restrict__ lo,
                                  const int* __restrict__ up, const int* __restrict__ diag,
                                  const double* __restrict__ val, T* __restrict__ inv, T* __restrict__ Lo,
                                  T* __restrict__ Up, T* __restrict__ WL, T* __restrict__ WR) {
  constexpr int BB = B * B;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int l = t / 8, sg = t % 8;
  if (l >= nl) return;
  const int L = len[l], ns = nseg[l];
  if (L == 0 || ns == 0 || sg >= ns) return;
  int a, b;
  line_segment(L, ns, sg, a, b);
  const int m = a + (b - a + 1) / 2;
  auto AT = [&](int k) -> size_t { return (size_t)k * nl + l; };
  auto vblk = [&](int q) { Blk<B> x; x.load(val + (size_t)q * BB); return x; };
  Blk<B> zero;
  zero.zero();
  // factors: top half [a, m) downwards, bottom half (m, b] upwards
  Blk<B> Pt, Ut, Pb, Lb;
  for (int k = a; k < m; ++k) {
    const size_t at = AT(k);
    Blk<B> Dk = vblk(diag[cell[at]]), A = zero;
    if (k > a) {
      A = vblk(lo[at]);
      Dk = bsub(Dk, bmul(bmul(A, Pt), Ut));
    }
    A.store(Lo + at * BB);
    Pt = binv(Dk);
    Pt.store(inv + at * BB);
    Ut = vblk(up[at]);
    bmul(Pt, Ut).store(Up + at * BB);
  }
  for (int k = b; k > m; --k) {
    const size_t at = AT(k);
    Blk<B> Dk = vblk(diag[cell[at]]), A = zero;
    if (k < b) {
      A = vblk(up[at]);
      Dk = bsub(Dk, bmul(bmul(A, Pb), Lb));
    }
    A.store(Lo + at * BB);
    Pb = binv(Dk);
    Pb.store(inv + at * BB);
    Lb = vblk(lo[at]);
    bmul(Pb, Lb).store(Up + at * BB);
  }
  {
    const size_t at = AT(m);
    Blk<B> Dk = vblk(diag[cell[at]]), A = zero, C = zero;
    if (m > a) {
      A = vblk(lo[at]);
      Dk = bsub(Dk, bmul(bmul(A, Pt), Ut));
    }
    if (m < b) {
      C = vblk(up[at]);
      Dk = bsub(Dk, bmul(bmul(C, Pb), Lb));
    }
    A.store(Lo + at * BB);
    C.store(Up + at * BB);
    binv(Dk).store(inv + at * BB);
  }
  // spikes: A_seg W = E_a A(a, a-1) (left) or E_b A(b, b+1) (right); the
  // forward results are kept in W until the backward walk overwrites them
  auto spike = [&](T* W, bool left) {
    const Blk<B> R = left ? vblk(lo[AT(a)]) : vblk(up[AT(b)]);
    Blk<B> yp = zero, fi, fl;
    for (int k = a; k < m; ++k) {
      const size_t at = AT(k);
      fi.load(inv + at * BB);
      fl.load(Lo + at * BB);
      const Blk<B> rk = (left && k == a) ? R : zero;
      yp = bmul(fi, bsub(rk, bmul(fl, yp)));
      yp.store(W + at * BB);
    }
    const Blk<B> ytop = yp;
    yp = zero;
    for (int k = b; k > m; --k) {
      const size_t at = AT(k);
      fi.load(inv + at * BB);
      fl.load(Lo + at * BB);
      const Blk<B> rk = (!left && k == b) ? R : zero;
      yp = bmul(fi, bsub(rk, bmul(fl, yp)));
      yp.store(W + at * BB);
    }
    const Blk<B> ybot = yp;
    Blk<B> ym;
    {
      const size_t at = AT(m);
      Blk<B> fu;
      fi.load(inv + at * BB);
      fl.load(Lo + at * BB);
      fu.load(Up + at * BB);
      const Blk<B> rm = ((left && m == a) || (!left && m == b)) ? R : zero;
      ym = bmul(fi, bsub(bsub(rm, bmul(fl, ytop)), bmul(fu, ybot)));
      ym.store(W + at * BB);
    }
    Blk<B> wn = ym, yk, pk;
    for (int k = m - 1; k >= a; --k) {
      const size_t at = AT(k);
      yk.load(W + at * BB);
      pk.load(Up + at * BB);
      wn = bsub(yk, bmul(pk, wn));
      wn.store(W + at * BB);
    }
    wn = ym;
    for (int k = m + 1; k <= b; ++k) {
      const size_t at = AT(k);
      yk.load(W + at * BB);
      pk.load(Up + at * BB);
      wn = bsub(yk, bmul(pk, wn));
      wn.store(W + at * BB);
    }
  };
  if (sg > 0) spike(WL, true);
  else
    for (int k = a; k <= b; ++k) zero.store(WL + AT(k) * BB);
  if (sg < ns - 1) spike(WR, false);
  else
    for (int k = a; k <= b; ++k) zero.store(WR + AT(k) * BB);
}

// The separators' Schur system of each line with nseg[l] >= 2 (one thread
// per line), factored by block Thomas: at separator k's position inv holds
// the inverse pivot D'_k^{-1}, Lo the coupling to separator k - 1, Up
// D'_k^{-1} times the coupling to separator k + 1, and WL / WR the
// separator's own couplings to the cells before and after it.
template <int B, class T>
__global__ void k_line_reduce(int nl, const int* __restrict__ len, const int* __restrict__ nseg,
                              const int* __restrict__ cell, const int* __restrict__ lo, const int* __restrict__ up,
                              const int* __restrict__ diag, const double* __restrict__ val, T* __restrict__ inv,
                              T* __restrict__ Lo, T* __restrict__ Up, T* __restrict__ WL, T* __restrict__ WR) {
  constexpr int BB = B * B;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nl) return;
  const int L = len[l], ns = nseg[l];
  if (L == 0 || ns < 2) return;
  auto AT = [&](int k) -> size_t { return (size_t)k * nl + l; };
  auto vblk = [&](int q) { Blk<B> x; x.load(val + (size_t)q * BB); return x; };
  Blk<B> Uprev;
  Uprev.zero();
  for (int k = 0; k < ns - 1; ++k) {
    const int sp = line_sep(L, ns, k);
    const size_t at = AT(sp), atp = AT(sp - 1), atn = AT(sp + 1);
    const Blk<B> Ass = vblk(diag[cell[at]]), Ls = vblk(lo[at]), Us = vblk(up[at]);
    Blk<B> wrp, wln, wlp, wrn;
    wrp.load(WR + atp * BB);
    wln.load(WL + atn * BB);
    wlp.load(WL + atp * BB);
    wrn.load(WR + atn * BB);
    const Blk<B> Tkk = bsub(bsub(Ass, bmul(Ls, wrp)), bmul(Us, wln));
    Blk<B> Tsub = bmul(Ls, wlp), Tsup = bmul(Us, wrn);
#pragma unroll
    for (int i = 0; i < B; ++i)
#pragma unroll
      for (int j = 0; j < B; ++j) {
        Tsub.v[i][j] = -Tsub.v[i][j];
        Tsup.v[i][j] = -Tsup.v[i][j];
      }
    const Blk<B> Dinv = binv(bsub(Tkk, bmul(Tsub, Uprev)));
    Uprev = bmul(Dinv, Tsup);
    Dinv.store(inv + at * BB);
    Tsub.store(Lo + at * BB);
    Uprev.store(Up + at * BB);
    Ls.store(WL + at * BB);
    Us.store(WR + at * BB);
  }
}

// One colour's partitioned line solves: a wave holds 32 / NS lines, lanes
// (line, segment) in both halves; lanes 0-31 walk their segments' top halves
// downwards, lanes 32-63 the bottom halves upwards, as k_line_gs_tw walks
// whole lines, keeping y and then g in LDS; one lane per line solves the
// separators' system from the segments' end values; every lane then
// corrects its cells by the spikes.  Lines shorter than the colour's longest
// may have fewer segments; their spare lanes idle.  Outputs as k_line_gs_tw:
// results by table position for the next colour, z by cell in the last
// sweep (Z).  The dynamic LDS holds (jcap + 1) x B x 64 + 32 x B doubles.
template <int B, class T, int D, int NS, bool Z>
__global__ __launch_bounds__(64) void k_line_gs_seg(int nl, int lc0, int jcap, const int* __restrict__ len,
                                                    const int* __restrict__ cell, const T* __restrict__ inv,
                                                    const T* __restrict__ Lo, const T* __restrict__ Up,
                                                    const T* __restrict__ WL, const T* __restrict__ WR,
                                                    const T* __restrict__ r2l, const T* __restrict__ xpl,
                                                    const PrecIO* __restrict__ io, T* __restrict__ ypos,
                                                    const double* __restrict__ pcw) {
  constexpr int BB = B * B, LPW = 32 / NS;
  extern __shared__ __align__(16) unsigned char line_lds[];
  double(*gs)[B][64] = reinterpret_cast<double(*)[B][64]>(line_lds);
  double(*ends)[64] = reinterpret_cast<double(*)[64]>(line_lds + sizeof(double) * (size_t)jcap * B * 64);
  double(*xs)[B] = reinterpret_cast<double(*)[B]>(line_lds + sizeof(double) * ((size_t)jcap + 1) * B * 64);
  gmem<double>* __restrict__ z = global_ptr(io->z);
  const int q = threadIdx.x, h = q >> 5, r = q & 31, li = r / NS, sg = r % NS;
  const int l = lc0 + blockIdx.x * LPW + li;
  if (len[lc0 + blockIdx.x * LPW] == 0) return;         // a wave of padding lines only (wave-uniform)
  const int L = len[l];
  const int ns = L > 0 ? line_nseg(L) : 1;
  int a = 0, b = -1;
  if (L > 0 && sg < ns) line_segment(L, ns, sg, a, b);
  const int n = b - a + 1;                              // 0 on an idle lane
  const int m = a + n / 2;
  const int J = n <= 0 ? 0 : (h == 0 ? n / 2 : n - 1 - n / 2);
  int Jw = J;
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) Jw = max(Jw, __shfl_xor(Jw, o));
  const int Jmax = (Jw + D - 1) / D * D;
  auto pos = [&](int j) -> int {
    const int jj = min(max(j, 0), max(J - 1, 0));
    return n <= 0 ? 0 : (h == 0 ? a + jj : b - jj);
  };
  // forward walk: y into LDS, rounded to T as the walks through memory round it
  T ri[D][BB], rl[D][BB], rr[D][B];
  auto fload = [&](int j, int u) {
    const size_t at = (size_t)pos(j) * nl + l;
#pragma unroll
    for (int k = 0; k < BB; ++k) {
      ri[u][k] = inv[at * BB + k];
      rl[u][k] = Lo[at * BB + k];
    }
#pragma unroll
    for (int c = 0; c < B; ++c) rr[u][c] = r2l[at * B + c];
  };
#pragma unroll
  for (int u = 0; u < D; ++u) fload(u, u);
  double y[B];
#pragma unroll
  for (int c = 0; c < B; ++c) y[c] = 0.0;
  for (int j0 = 0; j0 < Jmax; j0 += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 + u;
      double rhs[B], yn[B];
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double v = (double)rr[u][c];
#pragma unroll
        for (int k = 0; k < B; ++k) v -= (double)rl[u][c * B + k] * y[k];
        rhs[c] = v;
      }
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < B; ++k) v += (double)ri[u][c * B + k] * rhs[k];
        yn[c] = v;
      }
      if (j < J) {
#pragma unroll
        for (int c = 0; c < B; ++c) {
          y[c] = yn[c];
          gs[j][c][q] = (double)(T)yn[c];
        }
      }
      fload(j + D, u);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // the segment's middle cell
  double yo[B], xm[B];
#pragma unroll
  for (int c = 0; c < B; ++c) {
    yo[c] = __shfl_xor(y[c], 32);
    xm[c] = 0.0;
  }
  if (n > 0) {
    const size_t at = (size_t)m * nl + l;
    double rhs[B];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      double v = (double)r2l[at * B + c];
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const double yt = h == 0 ? y[k] : yo[k], yb = h == 0 ? yo[k] : y[k];
        v -= (double)Lo[at * BB + c * B + k] * yt + (double)Up[at * BB + c * B + k] * yb;
      }
      rhs[c] = v;
    }
#pragma unroll
    for (int c = 0; c < B; ++c) {
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < B; ++k) v += (double)inv[at * BB + c * B + k] * rhs[k];
      xm[c] = v;
    }
  }
  // backward walk: g = y - P g_next, into LDS
  T bp[D][BB];
  auto bload = [&](int j, int u) {
    const size_t at = (size_t)pos(j) * nl + l;
#pragma unroll
    for (int k = 0; k < BB; ++k) bp[u][k] = Up[at * BB + k];
  };
#pragma unroll
  for (int u = 0; u < D; ++u) bload(Jmax - 1 - u, u);
  double xn[B];
#pragma unroll
  for (int c = 0; c < B; ++c) xn[c] = xm[c];
  for (int j0 = Jmax - 1; j0 >= 0; j0 -= D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 - u;
      double tq[B];
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < B; ++k) v += (double)bp[u][c * B + k] * xn[k];
        tq[c] = v;
      }
      if (j < J) {
#pragma unroll
        for (int c = 0; c < B; ++c) {
          xn[c] = gs[j][c][q] - tq[c];
          gs[j][c][q] = xn[c];
        }
      }
      bload(j - D, u);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // end values: the top lane ends at the segment's first cell, the bottom
  // lane at its last (the middle when the half is empty)
#pragma unroll
  for (int c = 0; c < B; ++c) ends[c][q] = J > 0 ? xn[c] : xm[c];
  __syncthreads();
  double pc[B];
#pragma unroll
  for (int c = 0; c < B; ++c) pc[c] = pcw[c];
  // the separators of each line, one lane per line
  if (h == 0 && sg == 0 && L > 0 && ns > 1) {
    double zk[NS - 1][B], zp[B];
#pragma unroll
    for (int c = 0; c < B; ++c) zp[c] = 0.0;
#pragma unroll
    for (int k = 0; k < NS - 1; ++k) {
      if (k < ns - 1) {
        const size_t at = (size_t)line_sep(L, ns, k) * nl + l;
        const int ql = 32 + li * NS + k, qf = li * NS + k + 1;   // segment k's bottom lane, k+1's top lane
        double rhs[B];
#pragma unroll
        for (int c = 0; c < B; ++c) {
          double v = (double)r2l[at * B + c];
#pragma unroll
          for (int e = 0; e < B; ++e)
            v -= (double)WL[at * BB + c * B + e] * ends[e][ql] + (double)WR[at * BB + c * B + e] * ends[e][qf] +
                 (double)Lo[at * BB + c * B + e] * zp[e];
          rhs[c] = v;
        }
#pragma unroll
        for (int c = 0; c < B; ++c) {
          double v = 0.0;
#pragma unroll
          for (int e = 0; e < B; ++e) v += (double)inv[at * BB + c * B + e] * rhs[e];
          zk[k][c] = v;
        }
#pragma unroll
        for (int c = 0; c < B; ++c) zp[c] = zk[k][c];
      }
    }
#pragma unroll
    for (int c = 0; c < B; ++c) zp[c] = 0.0;
#pragma unroll
    for (int k = NS - 2; k >= 0; --k) {
      if (k < ns - 1) {
        const size_t at = (size_t)line_sep(L, ns, k) * nl + l;
        double x[B];
#pragma unroll
        for (int c = 0; c < B; ++c) {
          double v = zk[k][c];
#pragma unroll
          for (int e = 0; e < B; ++e) v -= (double)Up[at * BB + c * B + e] * zp[e];
          x[c] = v;
        }
#pragma unroll
        for (int c = 0; c < B; ++c) {
          zp[c] = x[c];
          xs[li * NS + k][c] = x[c];
          if (ypos != nullptr) ypos[at * B + c] = (T)x[c];
        }
        if (Z) {
          const int cc = cell[at];
          const double xp = (double)xpl[at];
          if (cc >= 0)
#pragma unroll
            for (int c = 0; c < B; ++c) z[(size_t)cc * B + c] = x[c] + pc[c] * xp;
        }
      }
    }
  }
  __syncthreads();
  // correction of the segment's cells by its separators
  if (n <= 0) return;
  const bool hl = sg > 0, hr = sg < ns - 1;
  double xl[B], xr[B];
#pragma unroll
  for (int c = 0; c < B; ++c) {
    xl[c] = hl ? xs[li * NS + sg - 1][c] : 0.0;
    xr[c] = hr ? xs[li * NS + sg][c] : 0.0;
  }
  auto finish = [&](size_t at, const double (&g)[B], const T (&wl)[BB], const T (&wr)[BB], T xp, int cc) {
    double x[B];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      double v = g[c];
#pragma unroll
      for (int e = 0; e < B; ++e) v -= (double)wl[c * B + e] * xl[e] + (double)wr[c * B + e] * xr[e];
      x[c] = v;
    }
    if (ypos != nullptr)
#pragma unroll
      for (int c = 0; c < B; ++c) ypos[at * B + c] = (T)x[c];
    if (Z && cc >= 0)
#pragma unroll
      for (int c = 0; c < B; ++c) z[(size_t)cc * B + c] = x[c] + pc[c] * (double)xp;
  };
  for (int j0 = 0; j0 < Jmax; j0 += D) {
    T wl[D][BB], wr[D][BB], ox[D];
    int oc[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const size_t at = (size_t)pos(j0 + u) * nl + l;
#pragma unroll
      for (int k = 0; k < BB; ++k) {           // zero on the sides without a separator
        wl[u][k] = WL[at * BB + k];
        wr[u][k] = WR[at * BB + k];
      }
      ox[u] = Z ? xpl[at] : (T)0;
      oc[u] = Z ? cell[at] : -1;
    }
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int j = j0 + u;
      if (j < J) {
        double g[B];
#pragma unroll
        for (int c = 0; c < B; ++c) g[c] = gs[j][c][q];
        finish((size_t)pos(j) * nl + l, g, wl[u], wr[u], ox[u], oc[u]);
      }
    }
  }
  if (h == 0) {
    const size_t at = (size_t)m * nl + l;
    T wl[BB], wr[BB];
#pragma unroll
    for (int k = 0; k < BB; ++k) {
      wl[k] = WL[at * BB + k];
      wr[k] = WR[at * BB + k];
    }
    finish(at, xm, wl, wr, Z ? xpl[at] : (T)0, Z ? cell[at] : -1);
  }
}

// ---------------------------------------------------------------------------
// Block ILU(0) local stage on the Jacobian's block pattern, rows scheduled
// by levels: a row's level is one past the highest level among its lower
// (for the factorization and the forward solve) or upper (backward solve)
// neighbours, so the rows of one level are independent.  One thread per row.
// ---------------------------------------------------------------------------
// IKJ elimination of the rows of one level, in place on F (FP64 copy of the
// blocks): L_ik = A_ik U_kk^{-1} for each lower block in column order, and
// A_ij -= L_ik U_kj for every j > k the row holds (zero fill); then the
// inverse of the row's pivot U_ii.
template <int B>
__global__ void k_ilu_factor_level(int nr, const int* __restrict__ rows, const int* __restrict__ rowptr,
                                   const int* __restrict__ col, const int* __restrict__ diag,
                                   double* __restrict__ F, double* __restrict__ Dinv) {
  constexpr int BB = B * B;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
  const int end = rowptr[i + 1], di = diag[i];
  for (int q = rowptr[i]; q < di; ++q) {
    const int k = col[q];
    double a[B][B], dk[B][B], l[B][B];
    blk_load<B>(F + (size_t)q * BB, a);
    blk_load<B>(Dinv + (size_t)k * BB, dk);
#pragma unroll
    for (int r = 0; r < B; ++r)
#pragma unroll
      for (int c = 0; c < B; ++c) {
        double v = 0.0;
#pragma unroll
        for (int m = 0; m < B; ++m) v += a[r][m] * dk[m][c];
        l[r][c] = v;
      }
    blk_store<B, double>(F + (size_t)q * BB, l);
    int qi = q + 1;
    for (int p = diag[k] + 1; p < rowptr[k + 1]; ++p) {
      const int j = col[p];
      while (qi < end && col[qi] < j) ++qi;
      if (qi >= end) break;
      if (col[qi] != j) continue;
      double u[B][B], x[B][B];
      blk_load<B>(F + (size_t)p * BB, u);
      blk_load<B>(F + (size_t)qi * BB, x);
#pragma unroll
      for (int r = 0; r < B; ++r)
#pragma unroll
        for (int c = 0; c < B; ++c) {
          double v = x[r][c];
#pragma unroll
          for (int m = 0; m < B; ++m) v -= l[r][m] * u[m][c];
          x[r][c] = v;
        }
      blk_store<B, double>(F + (size_t)qi * BB, x);
    }
  }
  double piv[B][B], inv[B][B];
  blk_load<B>(F + (size_t)di * BB, piv);
  small_inverse<B>(piv, inv);
  blk_store<B, double>(Dinv + (size_t)i * BB, inv);
}

template <class T>
__global__ void k_convert(size_t n, const double* __restrict__ a, T* __restrict__ b) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)a[i];
}

// Forward solve of one level, in place: y_i -= sum_{k < i} L_ik y_k.
template <int B, class T>
__global__ void k_ilu_lower_level(int nr, const int* __restrict__ rows, const int* __restrict__ rowptr,
                                  const int* __restrict__ col, const int* __restrict__ diag, const T* __restrict__ F,
                                  T* __restrict__ y) {
  constexpr int BB = B * B;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
  double v[B];
#pragma unroll
  for (int a = 0; a < B; ++a) v[a] = (double)y[(size_t)i * B + a];
  for (int q = rowptr[i]; q < diag[i]; ++q) {
    const int k = col[q];
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) v[a] -= (double)F[(size_t)q * BB + a * B + b] * (double)y[(size_t)k * B + b];
  }
#pragma unroll
  for (int a = 0; a < B; ++a) y[(size_t)i * B + a] = (T)v[a];
}

// Backward solve of one level, in place: x_i = U_ii^{-1} (y_i - sum_{j > i}
// U_ij x_j); also z_i = x_i + pc xp_i into the preconditioner output.
template <int B, class T>
__global__ void k_ilu_upper_level(int nr, const int* __restrict__ rows, const int* __restrict__ rowptr,
                                  const int* __restrict__ col, const int* __restrict__ diag, const T* __restrict__ F,
                                  const T* __restrict__ Dinv, T* __restrict__ y, const T* __restrict__ xp,
                                  const double* __restrict__ pc, const PrecIO* __restrict__ io) {
  constexpr int BB = B * B;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nr) return;
  const int i = rows[t];
  double v[B], x[B];
#pragma unroll
  for (int a = 0; a < B; ++a) v[a] = (double)y[(size_t)i * B + a];
  for (int q = diag[i] + 1; q < rowptr[i + 1]; ++q) {
    const int j = col[q];
#pragma unroll
    for (int a = 0; a < B; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) v[a] -= (double)F[(size_t)q * BB + a * B + b] * (double)y[(size_t)j * B + b];
  }
#pragma unroll
  for (int a = 0; a < B; ++a) {
    double s = 0.0;
#pragma unroll
    for (int b = 0; b < B; ++b) s += (double)Dinv[(size_t)i * BB + a * B + b] * v[b];
    x[a] = s;
  }
  gmem<double>* __restrict__ z = global_ptr(io->z);
  const double xpi = (double)xp[i];
#pragma unroll
  for (int a = 0; a < B; ++a) {
    y[(size_t)i * B + a] = (T)x[a];
    z[(size_t)i * B + a] = x[a] + pc[a] * xpi;
  }
}

// ---------------------------------------------------------------------------
// Aggregation AMG kernels (scalar CSR, value type T)
// ---------------------------------------------------------------------------
// Coarse values as sums of fine values over the gather map; with rr, each
// fine value is also stored at its gather-map position (the fine level's
// entries grouped by coarse row, for the folded residual-and-restriction).
template <class T>
__global__ void k_galerkin(int cnnz, const int* __restrict__ gptr, const int* __restrict__ gsrc,
                           const T* __restrict__ fine, T* __restrict__ coarse, T* __restrict__ rr) {
  const int K = blockIdx.x * blockDim.x + threadIdx.x;
  if (K >= cnnz) return;
  double s = 0.0;
  for (int m = gptr[K]; m < gptr[K + 1]; ++m) {
    const T v = fine[gsrc[m]];
    s += (double)v;
    if (rr) rr[m] = v;
  }
  coarse[K] = (T)s;
}

template <class T>
__global__ void k_diag_inv(int n, const int* __restrict__ diag, const T* __restrict__ val, double omega,
                           T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = (double)val[diag[i]];
  dinv[i] = (T)(d != 0.0 ? omega / d : 0.0);
}

template <class T>
__global__ void k_jacobi0(int n, const T* __restrict__ dinv, const T* __restrict__ b, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = dinv[i] * b[i];
}

// mode 0: out = x + dinv*(b - A x)   (Jacobi sweep)
// mode 1: out = b - A x             (residual)
template <class T, int LPR>
__global__ __launch_bounds__(256) void k_csr_sweep(int n, const int* __restrict__ rowptr,
                                                   const int* __restrict__ col, const T* __restrict__ val,
                                                   const T* __restrict__ dinv, const T* __restrict__ b,
                                                   const T* __restrict__ x, T* __restrict__ out, int mode) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  T acc = 0;
  if (row < n) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) acc = fma(val[k], x[col[k]], acc);
  }
  acc = wave_sum(acc, LPR);
  if (row < n && lane == 0) {
    const T res = b[row] - acc;
    out[row] = mode == 0 ? x[row] + dinv[row] * res : res;
  }
}

// The same sweep on the sliced ELL copy: one thread per row, every load of
// a wave coalesced and independent of the others.
template <class T>
__global__ __launch_bounds__(256) void k_sell_sweep(int n, const int* __restrict__ sl_ptr,
                                                    const int* __restrict__ sl_len, const int* __restrict__ sl_col,
                                                    const T* __restrict__ sv, const T* __restrict__ dinv,
                                                    const T* __restrict__ b, const T* __restrict__ x,
                                                    T* __restrict__ out, int mode) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int sl = row >> 6;
  const int base = sl_ptr[sl] + (row & 63), len = sl_len[sl];
  T acc = 0;
#pragma unroll 8
  for (int k = 0; k < len; ++k) {
    const int q = base + 64 * k;
    acc = fma(sv[q], x[sl_col[q]], acc);
  }
  const T res = b[row] - acc;
  out[row] = mode == 0 ? x[row] + dinv[row] * res : res;
}

// The ELL sweep with 16-bit column offsets (column = row + offset): 6 bytes
// per stored FP32 entry instead of 8.
template <class T>
__global__ __launch_bounds__(256) void k_sell_sweep16(int n, const int* __restrict__ sl_ptr,
                                                      const int* __restrict__ sl_len,
                                                      const int16_t* __restrict__ sl_off, const T* __restrict__ sv,
                                                      const T* __restrict__ dinv, const T* __restrict__ b,
                                                      const T* __restrict__ x, T* __restrict__ out, int mode) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  const int sl = row >> 6;
  const int base = sl_ptr[sl] + (row & 63), len = sl_len[sl];
  T acc = 0;
#pragma unroll 8
  for (int k = 0; k < len; ++k) {
    const int q = base + 64 * k;
    acc = fma(sv[q], x[row + (int)sl_off[q]], acc);
  }
  const T res = b[row] - acc;
  out[row] = mode == 0 ? x[row] + dinv[row] * res : res;
}

template <class T>
__global__ void k_to_sell(int nnz, const int* __restrict__ map, const T* __restrict__ v, T* __restrict__ sv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) sv[map[k]] = v[k];
}

template <class T>
__global__ void k_add_inplace(int n, const T* __restrict__ a, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += a[i];
}

template <class T>
__global__ void k_restrict_agg(int nc, const int* __restrict__ aptr, const int* __restrict__ amem,
                               const T* __restrict__ r, T* __restrict__ bc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nc) return;
  T s = 0;
  for (int m = aptr[I]; m < aptr[I + 1]; ++m) s += r[amem[m]];
  bc[I] = s;
}

template <class T>
__global__ void k_prolong_add(int n, const int* __restrict__ agg, const T* __restrict__ xc, T alpha,
                              T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += alpha * xc[agg[i]];
}

// Folded forms of a V-cycle level's launches (CSR, LPR lanes per row), for
// levels where a launch costs more than the gathers it saves:
// k_csr_jacobi2: the zero-guess Jacobi step and the first sweep,
//   x0 = D^-1 b, out = x0 + D^-1 (b - A x0), with x0 formed where read;
// k_csr_resid_restrict: the residual and its restriction,
//   bc_I = sum over the aggregate's rows i of (b_i - (A x)_i), over the
//   aggregate's entries through the Galerkin gather map;
// k_csr_prolong_sweep: the coarse correction and the first post-sweep,
//   y = x + alpha P xc formed where read, out = y + D^-1 (b - A y).
// Each computes what the launches it replaces compute (the first and the
// third in the same order at the same LPR; the second sums each
// aggregate's residuals in another order).
template <class T, int LPR>
__global__ __launch_bounds__(256) void k_csr_jacobi2(int n, const int* __restrict__ rowptr,
                                                     const int* __restrict__ col, const T* __restrict__ val,
                                                     const T* __restrict__ dinv, const T* __restrict__ b,
                                                     T* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  T acc = 0;
  if (row < n) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) {
      const int c = col[k];
      const T x0 = dinv[c] * b[c];
      acc = fma(val[k], x0, acc);
    }
  }
  acc = wave_sum(acc, LPR);
  if (row < n && lane == 0) {
    const T x0 = dinv[row] * b[row];
    const T res = b[row] - acc;
    out[row] = fma(dinv[row], res, x0);
  }
}

// The fine entries of coarse row I are one contiguous range of the Galerkin
// gather map, [gptr[crow[I]], gptr[crow[I+1]]) (every entry of every row of
// the aggregate), stored in that order with their columns (rr_col, rr_val),
// so the lanes of a coarse row read them coalesced and split them evenly,
// whatever the aggregate's size.
template <class T, int LPR>
__global__ __launch_bounds__(256) void k_csr_resid_restrict(int nc, const int* __restrict__ aptr,
                                                            const int* __restrict__ amem,
                                                            const int* __restrict__ crow,
                                                            const int* __restrict__ gptr,
                                                            const int* __restrict__ rr_col,
                                                            const T* __restrict__ rr_val, const T* __restrict__ b,
                                                            const T* __restrict__ x, T* __restrict__ bc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = t / LPR, lane = t % LPR;
  T acc = 0;
  if (I < nc) {
    for (int m = aptr[I] + lane; m < aptr[I + 1]; m += LPR) acc += b[amem[m]];
    const int e1 = gptr[crow[I + 1]];
    for (int e = gptr[crow[I]] + lane; e < e1; e += LPR) acc = fma(-rr_val[e], x[rr_col[e]], acc);
  }
  acc = wave_sum(acc, LPR);
  if (I < nc && lane == 0) bc[I] = acc;
}

template <class T, int LPR>
__global__ __launch_bounds__(256) void k_csr_prolong_sweep(int n, const int* __restrict__ rowptr,
                                                           const int* __restrict__ col, const T* __restrict__ val,
                                                           const T* __restrict__ dinv, const T* __restrict__ b,
                                                           const T* __restrict__ x, const int* __restrict__ agg,
                                                           const T* __restrict__ xc, T alpha,
                                                           T* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int row = t / LPR, lane = t % LPR;
  T acc = 0;
  if (row < n) {
    const int e = rowptr[row + 1];
    for (int k = rowptr[row] + lane; k < e; k += LPR) {
      const int c = col[k];
      T y = x[c];
      y += alpha * xc[agg[c]];
      acc = fma(val[k], y, acc);
    }
  }
  acc = wave_sum(acc, LPR);
  if (row < n && lane == 0) {
    T y = x[row];
    y += alpha * xc[agg[row]];
    const T res = b[row] - acc;
    out[row] = fma(dinv[row], res, y);
  }
}

__global__ void k_dense_fill(int n, const int* __restrict__ rowptr, const int* __restrict__ col,
                             const double* __restrict__ val64, const float* __restrict__ val32,
                             double* __restrict__ dense) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = 0; j < n; ++j) dense[(size_t)i * n + j] = 0.0;
  for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
    dense[(size_t)i * n + col[k]] += val64 ? val64[k] : (double)val32[k];
}

// In-place Gauss-Jordan inversion of an n x n matrix by one workgroup.  The
// coarse Galerkin matrices of the pressure system are diagonally dominant,
// so no pivoting is needed; a vanishing pivot is replaced by a tiny shift.
__global__ __launch_bounds__(1024) void k_gauss_jordan(int n, double* __restrict__ a) {
  __shared__ double pivrow_scale;
  for (int k = 0; k < n; ++k) {
    if (threadIdx.x == 0) {
      double p = a[(size_t)k * n + k];
      if (p == 0.0) p = 1e-300;
      pivrow_scale = 1.0 / p;
    }
    __syncthreads();
    const double d = pivrow_scale;
    // scale pivot row (column k becomes d)
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      if (j == k) continue;
      a[(size_t)k * n + j] *= d;
    }
    __syncthreads();
    // eliminate column k from the other rows: one wave per row, lanes along
    // the row (column k and row k are only read here)
    {
      const int lane = threadIdx.x & 63, nwave = blockDim.x >> 6;
      for (int i = threadIdx.x >> 6; i < n; i += nwave) {
        if (i == k) continue;
        const double aik = a[(size_t)i * n + k];
        for (int j = lane; j < n; j += 64)
          if (j != k) a[(size_t)i * n + j] -= aik * a[(size_t)k * n + j];
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      if (i == k) continue;
      a[(size_t)i * n + k] *= -d;
    }
    __syncthreads();
    if (threadIdx.x == 0) a[(size_t)k * n + k] = d;
    __syncthreads();
  }
}

template <class T>
__global__ void k_dense_solve(int n, const double* __restrict__ ainv, const T* __restrict__ b, T* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s += ainv[(size_t)i * n + j] * (double)b[j];
  x[i] = (T)s;
}

// The coarse end of the V-cycle in one workgroup: levels small enough that
// a launch per smoothing sweep would cost more than the sweep.  Level k of
// the description is the k-th of these levels; the last one is solved with
// the dense inverse.  Same cycle as the launched levels: x = D^-1 b, further
// Jacobi sweeps, residual restricted to the coarser level, coarse
// correction scaled by alpha, post-smoothing sweeps.
template <class T>
struct CoarseLevel {
  int n;
  const int *rowptr, *col, *agg, *agg_ptr, *agg_mem;
  const T *v, *d;
  T *b, *x, *y;
};

constexpr int kCoarseThreads = 1024;
constexpr int kMaxCoarseLevels = 16;

template <class T>
__device__ void cb_sweep(const CoarseLevel<T>& a, const T* __restrict__ x, T* __restrict__ y) {
  for (int i = threadIdx.x; i < a.n; i += blockDim.x) {
    double acc = 0.0;
    for (int k = a.rowptr[i]; k < a.rowptr[i + 1]; ++k) acc += (double)a.v[k] * (double)x[a.col[k]];
    y[i] = (T)((double)x[i] + (double)a.d[i] * ((double)a.b[i] - acc));
  }
}

template <class T>
__global__ __launch_bounds__(kCoarseThreads) void k_coarse_cycle(int nlev, const CoarseLevel<T>* __restrict__ L,
                                                                 int pre, int post, double alpha,
                                                                 const double* __restrict__ dense) {
  __shared__ T* xs[kMaxCoarseLevels];
  const int tid = threadIdx.x;
  for (int l = 0; l + 1 < nlev; ++l) {
    const CoarseLevel<T> a = L[l];
    T* x = a.x;
    T* y = a.y;
    for (int i = tid; i < a.n; i += blockDim.x) x[i] = a.d[i] * a.b[i];
    __syncthreads();
    for (int it = 1; it < pre; ++it) {
      cb_sweep(a, x, y);
      __syncthreads();
      T* t = x; x = y; y = t;
    }
    if (tid == 0) xs[l] = x;
    const CoarseLevel<T> c = L[l + 1];
    for (int I = tid; I < c.n; I += blockDim.x) {
      double sI = 0.0;
      for (int m = a.agg_ptr[I]; m < a.agg_ptr[I + 1]; ++m) {
        const int i = a.agg_mem[m];
        double acc = 0.0;
        for (int k = a.rowptr[i]; k < a.rowptr[i + 1]; ++k) acc += (double)a.v[k] * (double)x[a.col[k]];
        sI += (double)a.b[i] - acc;
      }
      c.b[I] = (T)sI;
    }
    __syncthreads();
  }
  {
    // one wave per row of the dense inverse, its lanes along the row
    // (coalesced; a row-per-thread walk is one dependent load per entry)
    const CoarseLevel<T> a = L[nlev - 1];
    const int lane = tid & 63, nwave = blockDim.x >> 6;
    for (int i = tid >> 6; i < a.n; i += nwave) {
      double sI = 0.0;
      for (int j = lane; j < a.n; j += 64) sI += dense[(size_t)i * a.n + j] * (double)a.b[j];
      sI = wave_sum(sI, 64);
      if (lane == 0) a.x[i] = (T)sI;
    }
    if (tid == 0) xs[nlev - 1] = a.x;
    __syncthreads();
  }
  for (int l = nlev - 2; l >= 0; --l) {
    const CoarseLevel<T> a = L[l];
    T* x = xs[l];
    T* y = x == a.x ? a.y : a.x;
    const T* xc = L[l + 1].x;   // level l+1 leaves its result in its own x
    for (int i = tid; i < a.n; i += blockDim.x) x[i] = (T)((double)x[i] + alpha * (double)xc[a.agg[i]]);
    __syncthreads();
    for (int it = 0; it < post; ++it) {
      cb_sweep(a, x, y);
      __syncthreads();
      T* t = x; x = y; y = t;
    }
    if (x != a.x) {
      for (int i = tid; i < a.n; i += blockDim.x) a.x[i] = x[i];
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// CprSolver
// ---------------------------------------------------------------------------
// Levels of at least kSellRows rows also keep their matrix in sliced ELL
// form for the smoother: slices of 64 rows (one wave), entry k of row r at
// sl_ptr[r / 64] + 64 k + r % 64, each slice padded to its longest row with
// zero values (column: the row itself).  The CSR values stay the source
// (Galerkin products read them); sl_map sends CSR entries to ELL slots.
// Measured: the 1.12M-row level of the oil-water field case gains (22.1 ->
// 23.8 Mcell-Newton-its/s over the run); levels of 140k-200k rows do not
// (their sweeps stay latency-bound), hence the threshold.
constexpr int kSellRows = 262144;
// with autotuning, levels from this size keep an ELL copy as a candidate
constexpr int kSellTuneRows = 65536;
struct AmgLevel {
  int n = 0, nnz = 0, ncoarse = 0;
  DArr<int> rowptr, col, diag, agg, agg_ptr, agg_mem, gal_ptr, gal_src;
  DArr<double> v64, d64, x64, y64, b64, r64;
  DArr<float> v32, d32, x32, y32, b32, r32;
  int nslice = 0;
  size_t sell_slots = 0;
  bool sell16 = false;                // sl_off holds every column as a 16-bit offset from its row
  DArr<int16_t> sl_off;
  DArr<int> sl_ptr, sl_len, sl_col, sl_map;
  DArr<double> sv64;
  DArr<float> sv32;
  // launch shapes of this level's V-cycle launches: lanes per CSR row, the
  // ELL copy (when the level has one) or CSR, and folded launches (CSR);
  // fixed defaults, or chosen by timing at the first setup when autotuned
  int lpr = 8;
  bool use_sell = true;
  bool fold = false;
  bool tuned = false;
  // for folding: the fine entries grouped by coarse row (gather-map order):
  // their columns (static) and values (per setup)
  DArr<int> rr_col;
  DArr<double> rr64;
  DArr<float> rr32;
};

static bool sweeps_from_sell(const AmgLevel& a) { return a.nslice > 0 && a.use_sell; }

// Candidate shapes of a level: 0 the ELL copy (CSR with 8 lanes where the
// level has none), 1-3 CSR with 4, 8, 16 lanes per row, 4-6 the same folded.
static void set_level_shape(AmgLevel& a, int cand) {
  a.use_sell = cand == 0;
  a.lpr = cand == 0 ? 8 : (4 << ((cand - 1) % 3));
  a.fold = cand >= 4;
}

struct SegFGMRES;
int cpr_solve_segmented(CprSolver* C, const BlockSystem& S, const double* b, double* x,
                        const LinearSolverOptions& o, SolveInfo& info, hipStream_t s);

class CprSolver {
 public:
  std::unique_ptr<SegFGMRES> seg;     // per-segment Krylov state (ensembles), when set
  ~CprSolver();
  std::vector<std::unique_ptr<AmgLevel>> L;
  bool built = false;
  int coarse_n = 0;
  DArr<double> dense;                 // coarsest: inverse (n x n)
  DArr<double> inv64, tail_inv, Jp64, Jps64;   // Jps: Jp in the finest level's ELL slot order
  DArr<float> inv32, Jp32, Jps32;
  DArr<double> hp, hmat;              // per-well scratch
  // Krylov workspace
  DArr<double> V, Z, w, partial, hdev, ydev, r;
  int m_alloc = 0;
  size_t n_alloc = 0, ld = 0;
  int prec_setup = 0;
  const double* wdev = nullptr;       // pressure weights of the current setup (row combination)
  const double* pcdev = nullptr;      // prolongation weights of the current setup (column spread)
  DArr<double> pc_unit;               // e_0: the pressure column
  EventTimer spmv_t, prec_t;
  // line-solve launches timed in eager applications (never inside a
  // capture), with the algorithmic bytes of the launches timed
  EventTimer lgs_t;
  double lgs_bytes = 0.0;
  bool capturing = false;
  std::vector<long long> color_cells;                // cells on the lines of each colour
  // partitioned lines: segments per line of each colour's longest line (1:
  // whole lines) and the longest segment half-walk
  std::vector<int> color_ns, color_jcap;
  bool seg_active = false;                           // this setup factored the partitioned lines
  DArr<int> ln_nseg;                                 // segments of each line (0: whole-line factors)
  // line local stage: position-major line tables and block-tridiagonal factors
  int nlines = 0, maxlen = 0, longest = 0, ncolor = 0;
  std::vector<int> color_ptr, color_len;             // line range and longest line of each colour
  std::vector<int> user_line_ptr, user_line_cells;   // lines given by the caller
  DArr<int> ln_len, ln_cell, ln_lo, ln_up;
  DArr<unsigned char> ln_color;                      // colour of each cell's line
  int noff = 0;                                      // off-line coupling planes
  DArr<int> ln_pos, off_q, off_nbr;
  DArr<unsigned char> off_col;
  DArr<double> lf_inv64, lf_lo64, lf_up64, r2_64, xpl64, yc64, r1l64, aoff64, lf_wl64, lf_wr64;
  DArr<float> lf_inv32, lf_lo32, lf_up32, r2_32, xpl32, yc32, r1l32, aoff32, lf_wl32, lf_wr32;
  // coarse levels handled by one workgroup: first such level, descriptions
  int coarse_l0 = -1;
  int dot2_elems = 0;                 // elements per thread of k_multidot2 (0: not chosen yet)
  size_t dot2_n = 0;                  // the system length it was chosen for
  DArr<char> coarse_desc;
  // captured preconditioner application
  PrecIO* io = nullptr;
  hipStream_t cap_s = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_ok = false;
  LinearSolverOptions graph_opts;

  // repeated pressure cycles: original right-hand side and accumulated solution
  DArr<double> pb64, pacc64;
  DArr<float> pb32, pacc32;
  // block ILU(0) local stage: rows in level order for the factorization and
  // forward solve (lower levels) and for the backward solve (upper levels)
  std::vector<int> ilu_lo_ptr, ilu_up_ptr;
  DArr<int> ilu_lo_rows, ilu_up_rows;
  DArr<double> ilu_w64, ilu_dw64;                    // FP64 factorization work: blocks, pivot inverses
  DArr<double> ilu_f64, ilu_di64, ilu_y64, ilu_xp64;
  DArr<float> ilu_f32, ilu_di32, ilu_y32, ilu_xp32;
  // a caller's pressure-stage solver (host callback) in place of the AMG
  PressureSolveFn pfn = nullptr;
  void* puser = nullptr;
  std::vector<double> pb_host, px_host;
};

void cpr_set_pressure_solver(CprSolver* C, PressureSolveFn fn, void* user) {
  C->pfn = fn;
  C->puser = user;
  C->graph_ok = false;
}

void cpr_profile(CprSolver* C, bool on) {
  C->spmv_t.on = on;
  C->prec_t.on = on;
  C->spmv_t.reset();
  C->prec_t.reset();
  C->lgs_t.on = on;
  C->lgs_t.reset();
  C->lgs_bytes = 0.0;
}

void cpr_profile_line_read(CprSolver* C, double* out) {
  int n = 0;
  out[1] = C->lgs_t.total_ms(&n);
  out[0] = n;
  out[2] = C->lgs_bytes;
}

int cpr_launch_shapes(CprSolver* C, const BlockSystem& S, int* out, int cap) {
  std::vector<int> v{S.spmv_lpr, 0};
  const int nL = (int)C->L.size();
  for (int l = 0; l + 1 < nL && l < C->coarse_l0; ++l) {
    const AmgLevel& a = *C->L[l];
    v.insert(v.end(), {a.n, sweeps_from_sell(a) ? 1 : 0, a.lpr, a.fold ? 1 : 0, a.tuned ? 1 : 0});
    ++v[1];
  }
  v.push_back(C->dot2_elems);
  const int n = std::min(cap, (int)v.size());
  std::copy(v.begin(), v.begin() + n, out);
  return n;
}

void cpr_profile_read(CprSolver* C, double* out) {
  int n1 = 0, n2 = 0;
  out[1] = C->spmv_t.total_ms(&n1);
  out[3] = C->prec_t.total_ms(&n2);
  out[0] = n1;
  out[2] = n2;
}

CprSolver* cpr_create() { return new CprSolver(); }
void cpr_destroy(CprSolver* c) { delete c; }
bool cpr_has_hierarchy(const CprSolver* c) { return c && c->built; }

void cpr_set_lines(CprSolver* C, int nlines, const int* ptr, const int* cells) {
  C->user_line_ptr.assign(ptr, ptr + nlines + 1);
  C->user_line_cells.assign(cells, cells + ptr[nlines]);
  C->built = false;   // rebuild the static structures with these lines
  C->graph_ok = false;
}

void cpr_build_hierarchy(CprSolver* C, const BlockSystem& S, const std::vector<double>& weights,
                         const LinearSolverOptions& o, hipStream_t s) {
  const auto t_build = std::chrono::steady_clock::now();
  // host copy of the BCSR pattern
  HostPattern fine;
  fine.n = S.nc;
  fine.rowptr.resize(S.nc + 1);
  fine.col.resize(S.nnzb);
  fine.diag.resize(S.nc);
  PMX_HIP(hipMemcpyAsync(fine.rowptr.data(), S.rowptr.p, sizeof(int) * (S.nc + 1), hipMemcpyDeviceToHost, s));
  PMX_HIP(hipMemcpyAsync(fine.col.data(), S.col.p, sizeof(int) * S.nnzb, hipMemcpyDeviceToHost, s));
  PMX_HIP(hipMemcpyAsync(fine.diag.data(), S.diag.p, sizeof(int) * S.nc, hipMemcpyDeviceToHost, s));
  PMX_HIP(hipStreamSynchronize(s));
  PMX_CHECK((int)weights.size() == S.nnzb, "hierarchy weights must have one entry per block");
  AggregationParams prm;
  prm.theta = o.theta;
  prm.coarse_size = o.coarse_size;
  prm.pairwise_passes = o.pairwise_passes;
  std::vector<HostLevel> H;
  build_aggregation(fine, weights, prm, H);
  C->L.clear();
  for (size_t l = 0; l < H.size(); ++l) {
    auto lv = std::make_unique<AmgLevel>();
    const HostLevel& h = H[l];
    lv->n = h.pat.n;
    lv->nnz = h.pat.rowptr[h.pat.n];
    lv->rowptr.upload(h.pat.rowptr, s);
    lv->col.upload(h.pat.col, s);
    lv->diag.upload(h.pat.diag, s);
    lv->ncoarse = h.ncoarse;
    if (l + 1 < H.size() && lv->n >= (o.autotune ? kSellTuneRows : kSellRows)) {
      const int n = lv->n, ns = (n + 63) / 64;
      const std::vector<int>& rp = h.pat.rowptr;
      std::vector<int> ptr(ns + 1), len(ns), map(lv->nnz);
      size_t off = 0;
      for (int q = 0; q < ns; ++q) {
        int L = 0;
        for (int r = 64 * q; r < std::min(n, 64 * q + 64); ++r) L = std::max(L, rp[r + 1] - rp[r]);
        ptr[q] = (int)off;
        len[q] = L;
        off += 64 * (size_t)L;
      }
      ptr[ns] = (int)off;
      std::vector<int> cl(off, 0);
      for (int q = 0; q < ns; ++q)
        for (int r = 64 * q; r < 64 * q + 64; ++r)
          for (int k = 0; k < len[q]; ++k) {
            const size_t slot = (size_t)ptr[q] + 64 * (size_t)k + (r - 64 * q);
            if (r < n && k < rp[r + 1] - rp[r]) {
              cl[slot] = h.pat.col[rp[r] + k];
              map[rp[r] + k] = (int)slot;
            } else {
              cl[slot] = std::min(r, n - 1);   // padding: zero value
            }
          }
      lv->nslice = ns;
      lv->sell_slots = off;
      lv->sl_ptr.upload(ptr, s);
      lv->sl_len.upload(len, s);
      lv->sl_col.upload(cl, s);
      // the same columns as 16-bit offsets from the row, when they all fit
      // (structured grids: +-1, +-nx, +-nx ny), for the sweeps
      std::vector<int16_t> d16(off, 0);
      bool fits = true;
      for (int q = 0; q < ns && fits; ++q)
        for (int r = 64 * q; r < std::min(n, 64 * q + 64) && fits; ++r)
          for (int k = 0; k < len[q]; ++k) {
            const size_t slot = (size_t)ptr[q] + 64 * (size_t)k + (r - 64 * q);
            const int dlt = cl[slot] - r;
            if (dlt < -32768 || dlt > 32767) {
              fits = false;
              break;
            }
            d16[slot] = (int16_t)dlt;
          }
      lv->sell16 = fits;
      if (fits) lv->sl_off.upload(d16, s);
      lv->sl_map.upload(map, s);
    }
    if (l + 1 < H.size()) {
      lv->agg.upload(h.agg, s);
      lv->agg_ptr.upload(h.agg_ptr, s);
      lv->agg_mem.upload(h.agg_mem, s);
      lv->gal_ptr.upload(h.gal_ptr, s);
      lv->gal_src.upload(h.gal_src, s);
      // columns of the fine entries in gather-map order (folded launches)
      if (o.autotune || o.level_shape >= 4) {
        std::vector<int> rc(h.gal_src.size());
        for (size_t e = 0; e < rc.size(); ++e) rc[e] = h.pat.col[h.gal_src[e]];
        lv->rr_col.upload(rc, s);
      }
    }
    C->L.push_back(std::move(lv));
  }
  // lines for the local stage, from the same strength weights
  HostLines HL;
  if (!C->user_line_ptr.empty())
    lines_from_list(fine, (int)C->user_line_ptr.size() - 1, C->user_line_ptr.data(), C->user_line_cells.data(), HL);
  else
    build_lines(fine, weights, HL);
  C->nlines = HL.nlines;
  C->maxlen = HL.maxlen;
  C->longest = HL.longest;
  C->ncolor = HL.ncolor;
  C->color_ptr = HL.color_ptr;
  C->color_len.assign(HL.ncolor, 0);
  C->color_cells.assign(HL.ncolor, 0);
  for (int c = 0; c < HL.ncolor; ++c)
    for (int l = HL.color_ptr[c]; l < HL.color_ptr[c + 1]; ++l) C->color_cells[c] += HL.len[l];
  for (int c = 0; c < HL.ncolor; ++c)
    if (HL.color_ptr[c + 1] > HL.color_ptr[c]) C->color_len[c] = HL.len[HL.color_ptr[c]];
  // partitioned lines: a colour whose longest line has 32 cells or more is
  // solved by segments, unless its longest half-walk would not fit the
  // 64 KiB of LDS a workgroup takes (3 x 3 blocks: 41 positions)
  {
    const int nc_ = HL.ncolor;
    C->color_ns.assign(nc_, 1);
    C->color_jcap.assign(nc_, 0);
    std::vector<int> nseg(HL.nlines, 0);
    for (int c = 0; c < nc_; ++c) {
      const int NS = line_nseg(C->color_len[c]);
      if (NS < 2) continue;
      int jcap = 0;
      for (int l = HL.color_ptr[c]; l < HL.color_ptr[c + 1]; ++l) {
        const int L = HL.len[l];
        if (L == 0) continue;
        const int ns = line_nseg(L);
        for (int k = 0; k < ns; ++k) {
          int a, b;
          line_segment(L, ns, k, a, b);
          jcap = std::max(jcap, (b - a + 1) / 2);
        }
      }
      if (jcap > 41) continue;
      C->color_ns[c] = NS;
      C->color_jcap[c] = std::max(jcap, 1);
      for (int l = HL.color_ptr[c]; l < HL.color_ptr[c + 1]; ++l) nseg[l] = HL.len[l] > 0 ? line_nseg(HL.len[l]) : 0;
    }
    C->ln_nseg.upload(nseg, s);
  }
  C->ln_len.upload(HL.len, s);
  C->ln_cell.upload(HL.cell, s);
  C->ln_lo.upload(HL.lo, s);
  C->ln_up.upload(HL.up, s);
  C->ln_color.upload(HL.cell_color, s);
  C->noff = HL.noff;
  C->ln_pos.upload(HL.pos, s);
  C->off_q.upload(HL.off_q, s);
  C->off_nbr.upload(HL.off_nbr, s);
  C->off_col.upload(HL.off_col, s);
  // ILU(0) level schedules of the block pattern
  {
    const int n = fine.n;
    std::vector<int> lev(n, 0);
    int nlo = 0, nup = 0;
    for (int i = 0; i < n; ++i) {
      int L = 0;
      for (int q = fine.rowptr[i]; q < fine.diag[i]; ++q) L = std::max(L, lev[fine.col[q]] + 1);
      lev[i] = L;
      nlo = std::max(nlo, L + 1);
    }
    auto bucket = [&](int nl, std::vector<int>& ptr, DArr<int>& rows) {
      ptr.assign(nl + 1, 0);
      for (int i = 0; i < n; ++i) ++ptr[lev[i] + 1];
      for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
      std::vector<int> r(n), at(ptr.begin(), ptr.end() - 1);
      for (int i = 0; i < n; ++i) r[at[lev[i]]++] = i;
      rows.upload(r, s);
    };
    bucket(nlo, C->ilu_lo_ptr, C->ilu_lo_rows);
    std::fill(lev.begin(), lev.end(), 0);
    for (int i = n - 1; i >= 0; --i) {
      int L = 0;
      for (int q = fine.diag[i] + 1; q < fine.rowptr[i + 1]; ++q) L = std::max(L, lev[fine.col[q]] + 1);
      lev[i] = L;
      nup = std::max(nup, L + 1);
    }
    bucket(nup, C->ilu_up_ptr, C->ilu_up_rows);
  }
  C->coarse_n = C->L.back()->n;
  PMX_CHECK(C->coarse_n <= 4096, "coarsest AMG level is too large for the dense solve");
  C->dense.alloc((size_t)C->coarse_n * C->coarse_n);
  PMX_HIP(hipStreamSynchronize(s));
  C->built = true;
  C->graph_ok = false;
  C->prec_setup = 0;
  if (o.verbose) {
    fprintf(stderr, "[pmx] hierarchy and lines built in %.1f ms\n",
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_build).count());
    fprintf(stderr, "[pmx] AMG hierarchy:");
    for (auto& lv : C->L) fprintf(stderr, " %d(%d)", lv->n, lv->nnz);
    fprintf(stderr, "\n[pmx] lines: %d in %d colours, longest %d\n", C->nlines, C->ncolor, C->longest);
  }
}


// Per-precision views of a level's arrays.
template <class T>
struct Lv;
template <>
struct Lv<double> {
  static DArr<double>& v(AmgLevel& l) { return l.v64; }
  static DArr<double>& sv(AmgLevel& l) { return l.sv64; }
  static DArr<double>& rr(AmgLevel& l) { return l.rr64; }
  static DArr<double>& d(AmgLevel& l) { return l.d64; }
  static DArr<double>& x(AmgLevel& l) { return l.x64; }
  static DArr<double>& y(AmgLevel& l) { return l.y64; }
  static DArr<double>& b(AmgLevel& l) { return l.b64; }
  static DArr<double>& r(AmgLevel& l) { return l.r64; }
  static DArr<double>& Jp(CprSolver& c) { return c.Jp64; }
  static DArr<double>& Jps(CprSolver& c) { return c.Jps64; }
  static DArr<double>& inv(CprSolver& c) { return c.inv64; }
  static DArr<double>& lf_inv(CprSolver& c) { return c.lf_inv64; }
  static DArr<double>& lf_lo(CprSolver& c) { return c.lf_lo64; }
  static DArr<double>& lf_up(CprSolver& c) { return c.lf_up64; }
  static DArr<double>& lf_wl(CprSolver& c) { return c.lf_wl64; }
  static DArr<double>& lf_wr(CprSolver& c) { return c.lf_wr64; }
  static DArr<double>& r2(CprSolver& c) { return c.r2_64; }
  static DArr<double>& xpl(CprSolver& c) { return c.xpl64; }
  static DArr<double>& yc(CprSolver& c) { return c.yc64; }
  static DArr<double>& r1l(CprSolver& c) { return c.r1l64; }
  static DArr<double>& aoff(CprSolver& c) { return c.aoff64; }
  static DArr<double>& ilu_f(CprSolver& c) { return c.ilu_f64; }
  static DArr<double>& ilu_di(CprSolver& c) { return c.ilu_di64; }
  static DArr<double>& ilu_y(CprSolver& c) { return c.ilu_y64; }
  static DArr<double>& ilu_xp(CprSolver& c) { return c.ilu_xp64; }
  static DArr<double>& pb(CprSolver& c) { return c.pb64; }
  static DArr<double>& pacc(CprSolver& c) { return c.pacc64; }
};
template <>
struct Lv<float> {
  static DArr<float>& v(AmgLevel& l) { return l.v32; }
  static DArr<float>& sv(AmgLevel& l) { return l.sv32; }
  static DArr<float>& rr(AmgLevel& l) { return l.rr32; }
  static DArr<float>& d(AmgLevel& l) { return l.d32; }
  static DArr<float>& x(AmgLevel& l) { return l.x32; }
  static DArr<float>& y(AmgLevel& l) { return l.y32; }
  static DArr<float>& b(AmgLevel& l) { return l.b32; }
  static DArr<float>& r(AmgLevel& l) { return l.r32; }
  static DArr<float>& Jp(CprSolver& c) { return c.Jp32; }
  static DArr<float>& Jps(CprSolver& c) { return c.Jps32; }
  static DArr<float>& inv(CprSolver& c) { return c.inv32; }
  static DArr<float>& lf_inv(CprSolver& c) { return c.lf_inv32; }
  static DArr<float>& lf_lo(CprSolver& c) { return c.lf_lo32; }
  static DArr<float>& lf_up(CprSolver& c) { return c.lf_up32; }
  static DArr<float>& lf_wl(CprSolver& c) { return c.lf_wl32; }
  static DArr<float>& lf_wr(CprSolver& c) { return c.lf_wr32; }
  static DArr<float>& r2(CprSolver& c) { return c.r2_32; }
  static DArr<float>& xpl(CprSolver& c) { return c.xpl32; }
  static DArr<float>& yc(CprSolver& c) { return c.yc32; }
  static DArr<float>& r1l(CprSolver& c) { return c.r1l32; }
  static DArr<float>& aoff(CprSolver& c) { return c.aoff32; }
  static DArr<float>& ilu_f(CprSolver& c) { return c.ilu_f32; }
  static DArr<float>& ilu_di(CprSolver& c) { return c.ilu_di32; }
  static DArr<float>& ilu_y(CprSolver& c) { return c.ilu_y32; }
  static DArr<float>& ilu_xp(CprSolver& c) { return c.ilu_xp32; }
  static DArr<float>& pb(CprSolver& c) { return c.pb32; }
  static DArr<float>& pacc(CprSolver& c) { return c.pacc32; }
};

template <class T>
static void tune_levels(CprSolver* C, const LinearSolverOptions& o, hipStream_t s);

template <int B, int M, int X, class T>
static void setup_t(CprSolver* C, const BlockSystem& S, const double* wdev, const LinearSolverOptions& o,
                    hipStream_t s) {
  PMX_CHECK(C->built, "AMG hierarchy not built");
  using V = Lv<T>;
  AmgLevel& f = *C->L[0];
  for (auto& lp : C->L) {
    AmgLevel& l = *lp;
    V::v(l).alloc(l.nnz);
    V::d(l).alloc(l.n);
    V::x(l).alloc(l.n);
    V::y(l).alloc(l.n);
    V::b(l).alloc(l.n);
    V::r(l).alloc(l.n);
  }
  DArr<T>& Jp = V::Jp(*C);
  DArr<T>& inv = V::inv(*C);
  Jp.alloc((size_t)S.nnzb * B);
  inv.alloc((size_t)S.nc * B * B);
  hipLaunchKernelGGL((k_pressure_matrix<B, T>), dim3(grid_for(S.nnzb)), dim3(256), 0, s, S.nnzb, S.val.p, wdev,
                     C->pcdev, V::v(f).p, Jp.p);
  launch_check();
  hipLaunchKernelGGL((k_block_inverse<B, T>), dim3(grid_for(S.nc)), dim3(256), 0, s, S.nc, S.diag.p, S.val.p,
                     inv.p);
  launch_check();
  if (o.local_stage == 1 && C->nlines > 0) {
    const size_t pos = (size_t)C->maxlen * C->nlines, m = pos * B * B;
    V::lf_inv(*C).alloc(m);
    V::lf_lo(*C).alloc(m);
    V::lf_up(*C).alloc(m);
    V::r2(*C).alloc(pos * B);
    V::xpl(*C).alloc(pos);
    V::yc(*C).alloc(pos * B);
    V::r1l(*C).alloc(pos * B);
    V::r1l(*C).zero(s);
    const size_t ne = (size_t)C->noff * pos;
    V::aoff(*C).alloc(std::max<size_t>(ne, 1) * B * B);
    if (ne > 0) {
      hipLaunchKernelGGL((k_gather_off<B, T>), dim3(grid_for(ne)), dim3(256), 0, s, ne, C->off_q.p, S.val.p,
                         V::aoff(*C).p);
      launch_check();
    }
    // entries past each line's end stay zero: the solves run over padded lengths
    V::lf_inv(*C).zero(s);
    V::lf_lo(*C).zero(s);
    V::lf_up(*C).zero(s);
    V::r2(*C).zero(s);
    V::xpl(*C).zero(s);
    C->seg_active = o.line_segments != 0 &&
                    std::any_of(C->color_ns.begin(), C->color_ns.end(), [](int v) { return v > 1; });
    hipLaunchKernelGGL((k_line_factor_tw<B, T>), dim3(grid_for(C->nlines, 64)), dim3(64), 0, s, C->nlines,
                       C->ln_len.p, C->seg_active ? C->ln_nseg.p : nullptr, C->ln_cell.p, C->ln_lo.p, C->ln_up.p,
                       S.diag.p, S.val.p, V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p);
    launch_check();
    if (C->seg_active) {
      V::lf_wl(*C).alloc(m);
      V::lf_wr(*C).alloc(m);
      hipLaunchKernelGGL((k_line_factor_seg<B, T>), dim3(grid_for((size_t)C->nlines * 8, 64)), dim3(64), 0, s,
                         C->nlines, C->ln_len.p, C->ln_nseg.p, C->ln_cell.p, C->ln_lo.p, C->ln_up.p, S.diag.p,
                         S.val.p, V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p, V::lf_wl(*C).p,
                         V::lf_wr(*C).p);
      launch_check();
      hipLaunchKernelGGL((k_line_reduce<B, T>), dim3(grid_for(C->nlines, 64)), dim3(64), 0, s, C->nlines,
                         C->ln_len.p, C->ln_nseg.p, C->ln_cell.p, C->ln_lo.p, C->ln_up.p, S.diag.p, S.val.p,
                         V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p, V::lf_wl(*C).p, V::lf_wr(*C).p);
      launch_check();
    }
  }
  if (o.local_stage == 2) {
    const size_t nb = (size_t)S.nnzb * B * B, nd = (size_t)S.nc * B * B;
    C->ilu_w64.alloc(nb);
    C->ilu_dw64.alloc(nd);
    PMX_HIP(hipMemcpyAsync(C->ilu_w64.p, S.val.p, sizeof(double) * nb, hipMemcpyDeviceToDevice, s));
    const int nlev = (int)C->ilu_lo_ptr.size() - 1;
    for (int l = 0; l < nlev; ++l) {
      const int r0 = C->ilu_lo_ptr[l], nr = C->ilu_lo_ptr[l + 1] - r0;
      hipLaunchKernelGGL((k_ilu_factor_level<B>), dim3(grid_for(nr, 64)), dim3(64), 0, s, nr, C->ilu_lo_rows.p + r0,
                         S.rowptr.p, S.col.p, S.diag.p, C->ilu_w64.p, C->ilu_dw64.p);
      launch_check();
    }
    V::ilu_f(*C).alloc(nb);
    V::ilu_di(*C).alloc(nd);
    V::ilu_y(*C).alloc((size_t)S.nc * B);
    V::ilu_xp(*C).alloc(S.nc);
    hipLaunchKernelGGL((k_convert<T>), dim3(grid_for(nb)), dim3(256), 0, s, nb, C->ilu_w64.p, V::ilu_f(*C).p);
    launch_check();
    hipLaunchKernelGGL((k_convert<T>), dim3(grid_for(nd)), dim3(256), 0, s, nd, C->ilu_dw64.p, V::ilu_di(*C).p);
    launch_check();
  }
  if (S.nw > 0) {
    C->tail_inv.alloc((size_t)S.nw * M * M);
    C->hp.alloc((size_t)S.nw * X);
    hipLaunchKernelGGL((k_tail_inverse<M>), dim3(grid_for(S.nw)), dim3(256), 0, s, S.nw, S.Dp.p, C->tail_inv.p);
    launch_check();
  }
  for (size_t l = 0; l < C->L.size(); ++l) {
    AmgLevel& a = *C->L[l];
    if (a.nslice > 0) {   // this level's values are final here: refresh the ELL copy
      V::sv(a).alloc(a.sell_slots);
      V::sv(a).zero(s);
      hipLaunchKernelGGL((k_to_sell<T>), dim3(grid_for(a.nnz)), dim3(256), 0, s, a.nnz, a.sl_map.p, V::v(a).p,
                         V::sv(a).p);
      launch_check();
      if (l == 0 && ((o.local_stage == 1 && C->nlines > 0) || o.local_stage == 2)) {
        PMX_CHECK(a.nnz == S.nnzb && a.n == S.nc, "the finest level shares the Jacobian's block pattern");
        V::Jps(*C).alloc(a.sell_slots * B);
        V::Jps(*C).zero(s);
        hipLaunchKernelGGL((k_to_sell_vec<T, B>), dim3(grid_for(a.nnz)), dim3(256), 0, s, a.nnz, a.sl_map.p,
                           V::Jp(*C).p, V::Jps(*C).p);
        launch_check();
      }
    }
    if (l + 1 < C->L.size()) {
      AmgLevel& c = *C->L[l + 1];
      // the fine values in gather-map order for the folded residual-and-
      // restriction, on levels that can fold (CSR sweeps)
      T* rr = nullptr;
      if (a.rr_col.n > 0) {
        V::rr(a).alloc(a.rr_col.n);
        rr = V::rr(a).p;
      }
      hipLaunchKernelGGL((k_galerkin<T>), dim3(grid_for(c.nnz)), dim3(256), 0, s, c.nnz, a.gal_ptr.p, a.gal_src.p,
                         V::v(a).p, V::v(c).p, rr);
      launch_check();
      hipLaunchKernelGGL((k_diag_inv<T>), dim3(grid_for(a.n)), dim3(256), 0, s, a.n, a.diag.p, V::v(a).p,
                         o.jacobi_omega, V::d(a).p);
      launch_check();
    } else {
      const int n = a.n;
      const double* v64 = nullptr;
      const float* v32 = nullptr;
      if constexpr (std::is_same<T, double>::value) v64 = V::v(a).p; else v32 = V::v(a).p;
      hipLaunchKernelGGL(k_dense_fill, dim3(grid_for(n)), dim3(256), 0, s, n, a.rowptr.p, a.col.p, v64, v32,
                         C->dense.p);
      launch_check();
      hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(1024), 0, s, n, C->dense.p);
      launch_check();
    }
  }
  if (o.amg_cycles > 1) {
    V::pb(*C).alloc(f.n);
    V::pacc(*C).alloc(f.n);
  }
  // the coarse end of the cycle: the longest tail of levels with at most
  // coarse_block_rows rows each
  const int nL = (int)C->L.size();
  int l0 = nL - 1;
  while (l0 > 0 && C->L[l0 - 1]->n <= o.coarse_block_rows) --l0;
  PMX_CHECK(nL - l0 <= kMaxCoarseLevels, "too many coarse levels");
  std::vector<CoarseLevel<T>> desc(nL - l0);
  for (int l = l0; l < nL; ++l) {
    AmgLevel& a = *C->L[l];
    CoarseLevel<T>& d = desc[l - l0];
    d.n = a.n;
    d.rowptr = a.rowptr.p;
    d.col = a.col.p;
    d.agg = a.agg.p;
    d.agg_ptr = a.agg_ptr.p;
    d.agg_mem = a.agg_mem.p;
    d.v = V::v(a).p;
    d.d = V::d(a).p;
    d.b = V::b(a).p;
    d.x = V::x(a).p;
    d.y = V::y(a).p;
  }
  C->coarse_desc.upload(reinterpret_cast<const char*>(desc.data()), sizeof(CoarseLevel<T>) * desc.size(), s);
  C->coarse_l0 = l0;
  S.autotune = o.autotune;
  if (o.level_shape >= 0) {
    PMX_CHECK(o.level_shape <= 6, "level_shape is -1 or a candidate 0-6");
    for (auto& lp : C->L) set_level_shape(*lp, o.level_shape);
  } else if (o.autotune) {
    if (!stream_capturing(s)) {
      const auto t_tune = std::chrono::steady_clock::now();
      tune_levels<T>(C, o, s);
      if (o.verbose)
        fprintf(stderr, "[pmx] level shapes tuned in %.1f ms\n",
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_tune).count());
    }
  } else {
    for (auto& lp : C->L) {
      lp->lpr = 8;
      lp->use_sell = true;
      lp->fold = false;
      lp->tuned = false;
    }
  }
  PMX_HIP(hipStreamSynchronize(s));
  C->prec_setup = sizeof(T) == 8 ? 64 : 32;
  C->graph_ok = false;
  C->wdev = wdev;
}

void cpr_setup(CprSolver* C, const BlockSystem& S, const double* w_dev, const LinearSolverOptions& o,
               hipStream_t s) {
  cpr_setup(C, S, w_dev, nullptr, o, s);
}

void cpr_setup(CprSolver* C, const BlockSystem& S, const double* w_dev, const double* pc_dev,
               const LinearSolverOptions& o, hipStream_t s) {
  if (pc_dev == nullptr) {
    if (C->pc_unit.n == 0) {
      const double e0[4] = {1.0, 0.0, 0.0, 0.0};
      C->pc_unit.upload(e0, 4, s);
    }
    pc_dev = C->pc_unit.p;
  }
  C->pcdev = pc_dev;
  if (o.precision == 32)
    PMX_DISPATCH(S, (setup_t<B_, M_, X_, float>(C, S, w_dev, o, s)));
  else
    PMX_DISPATCH(S, (setup_t<B_, M_, X_, double>(C, S, w_dev, o, s)));
}

// Calls f with the lanes-per-row constant for lpr (4, 8 or 16).
template <class F>
static void with_lpr(int lpr, F&& f) {
  if (lpr == 4) f(std::integral_constant<int, 4>{});
  else if (lpr == 16) f(std::integral_constant<int, 16>{});
  else f(std::integral_constant<int, 8>{});
}

// One Jacobi sweep (mode 0: out = x + D^-1 (b - A x)) or residual (mode 1:
// out = b - A x) on level a, from its ELL copy when it has one and uses it.
template <class T>
static void level_sweep(AmgLevel& a, const T* b, const T* x, T* out, int mode, hipStream_t s) {
  using V = Lv<T>;
  if (sweeps_from_sell(a) && a.sell16)
    hipLaunchKernelGGL((k_sell_sweep16<T>), dim3(grid_for(a.n)), dim3(256), 0, s, a.n, a.sl_ptr.p, a.sl_len.p,
                       a.sl_off.p, V::sv(a).p, V::d(a).p, b, x, out, mode);
  else if (sweeps_from_sell(a))
    hipLaunchKernelGGL((k_sell_sweep<T>), dim3(grid_for(a.n)), dim3(256), 0, s, a.n, a.sl_ptr.p, a.sl_len.p,
                       a.sl_col.p, V::sv(a).p, V::d(a).p, b, x, out, mode);
  else
    with_lpr(a.lpr, [&](auto c) {
      constexpr int LPR = decltype(c)::value;
      hipLaunchKernelGGL((k_csr_sweep<T, LPR>), dim3(grid_for((size_t)a.n * LPR)), dim3(256), 0, s, a.n,
                         a.rowptr.p, a.col.p, V::v(a).p, V::d(a).p, b, x, out, mode);
    });
  launch_check();
}

// The launches of level l on the way down (pre-smoothing from a zero guess,
// residual, restriction into level l+1's right-hand side) and on the way up
// (coarse correction, post-smoothing), in their folded forms when the level
// uses them.
template <class T>
static void level_down(CprSolver* C, size_t l, const LinearSolverOptions& o, hipStream_t s) {
  using V = Lv<T>;
  AmgLevel& a = *C->L[l];
  AmgLevel& c = *C->L[l + 1];
  const int n = a.n;
  const bool fold = a.fold && !sweeps_from_sell(a);
  int it = 1;
  if (fold && o.pre_smooth >= 2) {
    with_lpr(a.lpr, [&](auto k) {
      constexpr int LPR = decltype(k)::value;
      hipLaunchKernelGGL((k_csr_jacobi2<T, LPR>), dim3(grid_for((size_t)n * LPR)), dim3(256), 0, s, n,
                         a.rowptr.p, a.col.p, V::v(a).p, V::d(a).p, V::b(a).p, V::y(a).p);
    });
    launch_check();
    std::swap(V::x(a).p, V::y(a).p);
    it = 2;
  } else {
    hipLaunchKernelGGL((k_jacobi0<T>), dim3(grid_for(n)), dim3(256), 0, s, n, V::d(a).p, V::b(a).p, V::x(a).p);
    launch_check();
  }
  for (; it < o.pre_smooth; ++it) {
    level_sweep<T>(a, V::b(a).p, V::x(a).p, V::y(a).p, 0, s);
    std::swap(V::x(a).p, V::y(a).p);
  }
  if (fold && a.rr_col.n > 0 && V::rr(a).n == a.rr_col.n) {
    with_lpr(a.lpr, [&](auto k) {
      constexpr int LPR = decltype(k)::value;
      hipLaunchKernelGGL((k_csr_resid_restrict<T, LPR>), dim3(grid_for((size_t)c.n * LPR)), dim3(256), 0, s, c.n,
                         a.agg_ptr.p, a.agg_mem.p, c.rowptr.p, a.gal_ptr.p, a.rr_col.p, V::rr(a).p, V::b(a).p,
                         V::x(a).p, V::b(c).p);
    });
    launch_check();
  } else {
    level_sweep<T>(a, V::b(a).p, V::x(a).p, V::r(a).p, 1, s);
    hipLaunchKernelGGL((k_restrict_agg<T>), dim3(grid_for(c.n)), dim3(256), 0, s, c.n, a.agg_ptr.p, a.agg_mem.p,
                       V::r(a).p, V::b(c).p);
    launch_check();
  }
}

template <class T>
static void level_up(CprSolver* C, size_t l, const LinearSolverOptions& o, hipStream_t s) {
  using V = Lv<T>;
  AmgLevel& a = *C->L[l];
  AmgLevel& c = *C->L[l + 1];
  const int n = a.n;
  const bool fold = a.fold && !sweeps_from_sell(a);
  int it = 0;
  if (fold && o.post_smooth >= 1) {
    with_lpr(a.lpr, [&](auto k) {
      constexpr int LPR = decltype(k)::value;
      hipLaunchKernelGGL((k_csr_prolong_sweep<T, LPR>), dim3(grid_for((size_t)n * LPR)), dim3(256), 0, s, n,
                         a.rowptr.p, a.col.p, V::v(a).p, V::d(a).p, V::b(a).p, V::x(a).p, a.agg.p, V::x(c).p,
                         (T)o.coarse_alpha, V::y(a).p);
    });
    launch_check();
    std::swap(V::x(a).p, V::y(a).p);
    it = 1;
  } else {
    hipLaunchKernelGGL((k_prolong_add<T>), dim3(grid_for(n)), dim3(256), 0, s, n, a.agg.p, V::x(c).p,
                       (T)o.coarse_alpha, V::x(a).p);
    launch_check();
  }
  for (; it < o.post_smooth; ++it) {
    level_sweep<T>(a, V::b(a).p, V::x(a).p, V::y(a).p, 0, s);
    std::swap(V::x(a).p, V::y(a).p);
  }
}

// One V-cycle on level l: x_l = V(b_l) (zero initial guess).
template <class T>
static void vcycle(CprSolver* C, size_t l, const LinearSolverOptions& o, hipStream_t s) {
  using V = Lv<T>;
  AmgLevel& a = *C->L[l];
  const int n = a.n;
  if ((int)l == C->coarse_l0) {
    hipLaunchKernelGGL((k_coarse_cycle<T>), dim3(1), dim3(kCoarseThreads), 0, s, (int)C->L.size() - (int)l,
                       reinterpret_cast<const CoarseLevel<T>*>(C->coarse_desc.p), o.pre_smooth, o.post_smooth,
                       o.coarse_alpha, C->dense.p);
    launch_check();
    return;
  }
  if (l + 1 == C->L.size()) {
    hipLaunchKernelGGL((k_dense_solve<T>), dim3(grid_for(n)), dim3(256), 0, s, n, C->dense.p, V::b(a).p,
                       V::x(a).p);
    launch_check();
    return;
  }
  level_down<T>(C, l, o, s);
  vcycle<T>(C, l + 1, o, s);
  level_up<T>(C, l, o, s);
}

// Autotuning of the levels above the coarse cycle: each level's down and up
// launches (with level l+1's vectors as they stand) timed for every shape
// (the ELL copy when the level has one; CSR with 4, 8 or 16 lanes per row,
// unfolded and folded) and the fastest kept.  The vectors are set to zero
// first (the timings do not depend on the values) and are overwritten by
// the first application.  Choices are kept per size class.
template <class T>
static void tune_levels(CprSolver* C, const LinearSolverOptions& o, hipStream_t s) {
  using V = Lv<T>;
  const int nL = (int)C->L.size();
  for (int l = 0; l + 1 < nL && l < C->coarse_l0; ++l) {
    AmgLevel& a = *C->L[l];
    if (a.tuned) continue;
    AmgLevel& c = *C->L[l + 1];
    const TuneKey k{2, (int)sizeof(T), octave4(a.n), octave4((double)a.nnz / std::max(a.n, 1)),
                    (a.nslice > 0 ? 1 : 0) + 2 * o.pre_smooth + 16 * o.post_smooth};
    int choice = -1;
    if (!tune_lookup(k, &choice)) {
      for (DArr<T>* v : {&V::b(a), &V::x(a), &V::y(a), &V::r(a), &V::b(c), &V::x(c)})
        PMX_HIP(hipMemsetAsync(v->p, 0, sizeof(T) * v->n, s));
      // candidate 0 (ELL) only where the level has the copy; the default is
      // ELL there and CSR with 8 lanes elsewhere
      std::vector<float> t(7, std::numeric_limits<float>::max());
      for (int cand = (a.nslice > 0 ? 0 : 1); cand < 7; ++cand) {
        set_level_shape(a, cand);
        t[cand] = time_launches(s, 5, [&] {
          level_down<T>(C, l, o, s);
          level_up<T>(C, l, o, s);
        });
      }
      choice = tune_pick(t, a.nslice > 0 ? 0 : 2);
      tune_store(k, choice);
    }
    set_level_shape(a, choice);
    a.tuned = true;
  }
}

template <int B, int M, int X, class T>
static void stage2(CprSolver* C, const BlockSystem& S, const LinearSolverOptions& o, hipStream_t s);

template <int B, int M, int X, class T>
static void apply_t(CprSolver* C, const BlockSystem& S, const double* wdev, const LinearSolverOptions& o,
                    hipStream_t s) {
  using V = Lv<T>;
  AmgLevel& f = *C->L[0];
  constexpr int LPR = 8;
  const size_t ncB = (size_t)S.nc * B;
  // stage 1: pressure
  hipLaunchKernelGGL((k_weight_restrict<B, T>), dim3(grid_for(S.nc)), dim3(256), 0, s, S.nc, wdev, C->io, V::b(f).p);
  launch_check();
  if (C->pfn != nullptr) {
    // the caller's pressure solver on the host: b = W r out, x back in
    const size_t n = f.n;
    std::vector<T> tb(n);
    PMX_HIP(hipMemcpyAsync(tb.data(), V::b(f).p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    PMX_HIP(hipStreamSynchronize(s));
    C->pb_host.assign(tb.begin(), tb.end());
    C->px_host.assign(n, 0.0);
    const int rc = C->pfn(C->puser, (int)n, C->pb_host.data(), C->px_host.data());
    PMX_CHECK(rc == 0, "the pressure-stage solver reported a failure");
    for (size_t i = 0; i < n; ++i) tb[i] = (T)C->px_host[i];
    PMX_HIP(hipMemcpyAsync(V::x(f).p, tb.data(), sizeof(T) * n, hipMemcpyHostToDevice, s));
    PMX_HIP(hipStreamSynchronize(s));
    stage2<B, M, X, T>(C, S, o, s);
    return;
  }
  vcycle<T>(C, 0, o, s);
  if (o.amg_cycles > 1) {
    // stationary iteration with the V-cycle: x <- x + V(b - A_p x)
    DArr<T>& b0 = V::pb(*C);
    DArr<T>& acc = V::pacc(*C);
    PMX_CHECK(b0.n == (size_t)f.n && acc.n == (size_t)f.n, "pressure-cycle buffers are set up with the solver");
    hipLaunchKernelGGL((k_copy<T>), dim3(grid_for(f.n)), dim3(256), 0, s, (size_t)f.n, V::b(f).p, b0.p);
    launch_check();
    for (int cyc = 1; cyc < o.amg_cycles; ++cyc) {
      std::swap(V::x(f).p, acc.p);
      level_sweep<T>(f, b0.p, acc.p, V::b(f).p, 1, s);
      vcycle<T>(C, 0, o, s);
      hipLaunchKernelGGL((k_add_inplace<T>), dim3(grid_for(f.n)), dim3(256), 0, s, f.n, acc.p, V::x(f).p);
      launch_check();
    }
  }
  stage2<B, M, X, T>(C, S, o, s);
}

// Stage 2: z = P xp + local(r - J P xp) with the local stage chosen by the
// options (cell block-Jacobi, or line block-tridiagonal solves); the well
// tail is solved exactly per well.
template <int B, int M, int X, class T>
static void stage2(CprSolver* C, const BlockSystem& S, const LinearSolverOptions& o, hipStream_t s) {
  using V = Lv<T>;
  AmgLevel& f = *C->L[0];
  constexpr int LPR = 8;
  const size_t ncB = (size_t)S.nc * B;
  if (S.nw > 0) {
    hipLaunchKernelGGL((k_cpr_wells<B, M, X, T>), dim3(S.nw), dim3(64), 0, s, S.nw, S.well_ptr.p, S.perf_cell.p,
                       S.Cp.p, S.V.p, C->tail_inv.p, V::x(f).p, C->io, ncB, C->hp.p, C->pcdev);
    launch_check();
  }
  if (o.local_stage == 2) {
    // r1 = r - J P xp (and the wells' mixture term) in cell order, then the
    // ILU(0) forward and backward solves level by level
    const bool mix = S.nw > 0 && S.has_mix && S.npc > 0;
    if (f.nslice > 0)
      hipLaunchKernelGGL((k_line_r1_sell<B, X, T>), dim3(grid_for(S.nc)), dim3(256), 0, s, S.nc, f.sl_ptr.p,
                         f.sl_len.p, f.sl_col.p, V::Jps(*C).p, V::x(f).p, C->io, mix ? S.pc_of.p : nullptr,
                         S.pc_ptr.p, S.pc_perf.p, S.perf_well.p, S.Um.p, C->hp.p, nullptr, V::ilu_y(*C).p,
                         V::ilu_xp(*C).p);
    else
      hipLaunchKernelGGL((k_line_r1<B, X, T, LPR>), dim3(grid_for((size_t)S.nc * LPR)), dim3(256), 0, s, S.nc,
                         S.rowptr.p, S.col.p, V::Jp(*C).p, V::x(f).p, C->io, mix ? S.pc_of.p : nullptr, S.pc_ptr.p,
                         S.pc_perf.p, S.perf_well.p, S.Um.p, C->hp.p, nullptr, V::ilu_y(*C).p, V::ilu_xp(*C).p);
    launch_check();
    const int nlo = (int)C->ilu_lo_ptr.size() - 1, nup = (int)C->ilu_up_ptr.size() - 1;
    for (int l = 0; l < nlo; ++l) {
      const int r0 = C->ilu_lo_ptr[l], nr = C->ilu_lo_ptr[l + 1] - r0;
      hipLaunchKernelGGL((k_ilu_lower_level<B, T>), dim3(grid_for(nr, 64)), dim3(64), 0, s, nr,
                         C->ilu_lo_rows.p + r0, S.rowptr.p, S.col.p, S.diag.p, V::ilu_f(*C).p, V::ilu_y(*C).p);
      launch_check();
    }
    for (int l = 0; l < nup; ++l) {
      const int r0 = C->ilu_up_ptr[l], nr = C->ilu_up_ptr[l + 1] - r0;
      hipLaunchKernelGGL((k_ilu_upper_level<B, T>), dim3(grid_for(nr, 64)), dim3(64), 0, s, nr,
                         C->ilu_up_rows.p + r0, S.rowptr.p, S.col.p, S.diag.p, V::ilu_f(*C).p, V::ilu_di(*C).p,
                         V::ilu_y(*C).p, V::ilu_xp(*C).p, C->pcdev, C->io);
      launch_check();
    }
    return;
  }
  if (o.local_stage == 1 && C->nlines > 0) {
    const bool mix = S.nw > 0 && S.has_mix && S.npc > 0;
    const int sweeps = std::max(1, o.local_sweeps);
    constexpr int D = line_depth<B, T>();
    const size_t npos = (size_t)C->maxlen * C->nlines;
    if (f.nslice > 0)
      hipLaunchKernelGGL((k_line_r1_sell<B, X, T>), dim3(grid_for(S.nc)), dim3(256), 0, s, S.nc, f.sl_ptr.p,
                         f.sl_len.p, f.sl_col.p, V::Jps(*C).p, V::x(f).p, C->io, mix ? S.pc_of.p : nullptr,
                         S.pc_ptr.p, S.pc_perf.p, S.perf_well.p, S.Um.p, C->hp.p, C->ln_pos.p, V::r1l(*C).p,
                         V::xpl(*C).p);
    else
      hipLaunchKernelGGL((k_line_r1<B, X, T, LPR>), dim3(grid_for((size_t)S.nc * LPR)), dim3(256), 0, s, S.nc,
                         S.rowptr.p, S.col.p, V::Jp(*C).p, V::x(f).p, C->io, mix ? S.pc_of.p : nullptr, S.pc_ptr.p,
                         S.pc_perf.p, S.perf_well.p, S.Um.p, C->hp.p, C->ln_pos.p, V::r1l(*C).p, V::xpl(*C).p);
    launch_check();
    for (int sw = 0; sw < sweeps; ++sw) {
      for (int c = 0; c < C->ncolor; ++c) {
        const int lc0 = C->color_ptr[c], nlc = C->color_ptr[c + 1] - lc0, Lc = C->color_len[c];
        if (nlc == 0) continue;
        PMX_CHECK(nlc % LINE_GROUP == 0, "line colours must fill whole waves");
        const bool last = sw + 1 == sweeps && c + 1 == C->ncolor;
        hipLaunchKernelGGL((k_line_rhs<B, T>), dim3(grid_for((size_t)nlc * Lc)), dim3(256), 0, s, C->nlines, lc0,
                           nlc, Lc, c, sw == 0 ? 1 : 0, C->noff, npos, C->ln_len.p, C->off_nbr.p, C->off_col.p,
                           V::aoff(*C).p, V::yc(*C).p, V::r1l(*C).p, V::r2(*C).p);
        launch_check();
        const bool zs = sw + 1 == sweeps;               // the last sweep writes the output
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool timed = !C->capturing && C->lgs_t.pair(&e0, &e1);
        const int NS = C->seg_active ? C->color_ns[c] : 1;
        if (NS > 1) {
          const int jcap = C->color_jcap[c];
          const size_t lds = sizeof(double) * ((size_t)(jcap + 1) * B * 64 + 32 * B);
          decltype(&k_line_gs_seg<B, T, D, 2, true>) sk;
          if (NS == 2) sk = zs ? k_line_gs_seg<B, T, D, 2, true> : k_line_gs_seg<B, T, D, 2, false>;
          else if (NS == 4) sk = zs ? k_line_gs_seg<B, T, D, 4, true> : k_line_gs_seg<B, T, D, 4, false>;
          else sk = zs ? k_line_gs_seg<B, T, D, 8, true> : k_line_gs_seg<B, T, D, 8, false>;
          const dim3 grid(nlc / 32 * NS);
          if (timed)
            hipExtLaunchKernelGGL(sk, grid, dim3(64), lds, s, e0, e1, 0, C->nlines, lc0, jcap, C->ln_len.p,
                                  C->ln_cell.p, V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p, V::lf_wl(*C).p,
                                  V::lf_wr(*C).p, V::r2(*C).p, V::xpl(*C).p, C->io, last ? nullptr : V::yc(*C).p,
                                  C->pcdev);
          else
            hipLaunchKernelGGL(sk, grid, dim3(64), lds, s, C->nlines, lc0, jcap, C->ln_len.p, C->ln_cell.p,
                               V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p, V::lf_wl(*C).p, V::lf_wr(*C).p,
                               V::r2(*C).p, V::xpl(*C).p, C->io, last ? nullptr : V::yc(*C).p, C->pcdev);
          launch_check();
          if (timed) {
            // per line cell five blocks and r read (a segment cell: D'^-1, L
            // and r going forward, P going back and its two spikes, zero at a
            // line's ends; a separator: its two couplings and three Thomas
            // blocks), the result written by table position unless this is
            // the last launch, and in the last sweep xp and the cell index
            // read and z written
            const double sv = sizeof(T), bb = B * B * sv;
            C->lgs_bytes += (double)C->color_cells[c] *
                            (5.0 * bb + B * sv + (last ? 0.0 : B * sv) + (zs ? sv + 4.0 + 8.0 * B : 0.0));
          }
          continue;
        }
        auto kern = zs ? k_line_gs_tw<B, T, D, true> : k_line_gs_tw<B, T, D, false>;
        if (timed)
          hipExtLaunchKernelGGL(kern, dim3(nlc / 32), dim3(64), 0, s, e0, e1, 0, C->nlines, lc0, C->ln_len.p,
                                C->ln_cell.p, V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p, V::r2(*C).p,
                                V::xpl(*C).p, C->io, last ? nullptr : V::yc(*C).p, C->pcdev);
        else
          hipLaunchKernelGGL(kern, dim3(nlc / 32), dim3(64), 0, s, C->nlines, lc0, C->ln_len.p, C->ln_cell.p,
                             V::lf_inv(*C).p, V::lf_lo(*C).p, V::lf_up(*C).p, V::r2(*C).p, V::xpl(*C).p, C->io,
                             last ? nullptr : V::yc(*C).p, C->pcdev);
        launch_check();
        if (timed) {
          // per line cell: forward D'^-1, L, r read and y written; backward
          // D'^-1, U, y read, and the result written by table position
          // unless this is the last launch; in the last sweep also the
          // pressure xp and the cell index read and z (FP64) written
          const double sv = sizeof(T);
          const double per_cell = 4.0 * B * B * sv + 3.0 * B * sv + (last ? 0.0 : B * sv) +
                                  (zs ? sv + 4.0 + 8.0 * B : 0.0);
          C->lgs_bytes += per_cell * (double)C->color_cells[c];
        }
      }
    }
    return;
  }
  hipLaunchKernelGGL((k_cpr_stage2<B, T, LPR>), dim3(grid_for((size_t)S.nc * LPR)), dim3(256), 0, s, S.nc,
                     S.rowptr.p, S.col.p, V::Jp(*C).p, V::x(f).p, C->io, V::inv(*C).p, C->pcdev);
  launch_check();
  if (S.nw > 0 && S.has_mix && S.npc > 0) {
    hipLaunchKernelGGL((k_cpr_mixfix<B, X, T>), dim3(grid_for(S.npc)), dim3(256), 0, s, S.npc, S.pc_cell.p,
                       S.pc_ptr.p, S.pc_perf.p, S.perf_well.p, S.Um.p, C->hp.p, V::inv(*C).p, C->io);
    launch_check();
  }
}

// Preconditioner without the pressure stage: block-Jacobi + tail only.
template <int B, int M, int X, class T>
static void apply_local_t(CprSolver* C, const BlockSystem& S, const LinearSolverOptions& o, hipStream_t s) {
  using V = Lv<T>;
  AmgLevel& f = *C->L[0];
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(f.n)), dim3(256), 0, s, (size_t)f.n, (T)0, V::x(f).p);
  launch_check();
  stage2<B, M, X, T>(C, S, o, s);
}

// The kernels of one application, reading r and z through C->io.
static void apply_sequence(CprSolver* C, const BlockSystem& S, const LinearSolverOptions& o, hipStream_t s) {
  if (o.precision == 32) {
    if (o.use_pressure_stage) PMX_DISPATCH(S, (apply_t<B_, M_, X_, float>(C, S, C->wdev, o, s)));
    else PMX_DISPATCH(S, (apply_local_t<B_, M_, X_, float>(C, S, o, s)));
  } else {
    if (o.use_pressure_stage) PMX_DISPATCH(S, (apply_t<B_, M_, X_, double>(C, S, C->wdev, o, s)));
    else PMX_DISPATCH(S, (apply_local_t<B_, M_, X_, double>(C, S, o, s)));
  }
}

static bool same_apply_options(const LinearSolverOptions& a, const LinearSolverOptions& b) {
  return a.precision == b.precision && a.use_pressure_stage == b.use_pressure_stage &&
         a.pre_smooth == b.pre_smooth && a.post_smooth == b.post_smooth && a.coarse_alpha == b.coarse_alpha &&
         a.amg_cycles == b.amg_cycles && a.local_stage == b.local_stage && a.local_sweeps == b.local_sweeps &&
         a.coarse_block_rows == b.coarse_block_rows && a.line_segments == b.line_segments;
}

// Capture the application once per setup (on a private stream: nothing runs
// while capturing) and replay it; a replay costs one launch instead of one
// per kernel.
static void capture_apply(CprSolver* C, const BlockSystem& S, const LinearSolverOptions& o) {
  if (C->cap_s == nullptr) PMX_HIP(hipStreamCreateWithFlags(&C->cap_s, hipStreamNonBlocking));
  if (C->gexec != nullptr) {
    PMX_HIP(hipGraphExecDestroy(C->gexec));
    C->gexec = nullptr;
  }
  PMX_HIP(hipStreamBeginCapture(C->cap_s, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  C->capturing = true;
  try {
    apply_sequence(C, S, o, C->cap_s);
    C->capturing = false;
  } catch (...) {
    C->capturing = false;
    (void)hipStreamEndCapture(C->cap_s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  PMX_HIP(hipStreamEndCapture(C->cap_s, &g));
  const hipError_t e = hipGraphInstantiate(&C->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  PMX_HIP(e);
  C->graph_opts = o;
  C->graph_ok = true;
}

void cpr_apply(CprSolver* C, const BlockSystem& S, const double* r, double* z, const LinearSolverOptions& o,
               hipStream_t s) {
  PMX_CHECK(C->prec_setup == o.precision, "preconditioner precision differs from its setup");
  PMX_CHECK(C->wdev != nullptr, "preconditioner weights unset");
  if (C->io == nullptr) PMX_HIP(hipMalloc(&C->io, sizeof(PrecIO)));
  hipLaunchKernelGGL(k_set_io, dim3(1), dim3(1), 0, s, C->io, r, z);
  launch_check();
  if (!o.use_graphs || C->pfn != nullptr) {   // a host pressure solver cannot be replayed from a graph
    apply_sequence(C, S, o, s);
    return;
  }
  if (!C->graph_ok || !same_apply_options(C->graph_opts, o)) capture_apply(C, S, o);
  PMX_HIP(hipGraphLaunch(C->gexec, s));
}


// ---------------------------------------------------------------------------
// Flexible GMRES (right preconditioning), FP64 Arnoldi with classical
// Gram-Schmidt (optionally twice).  Converged when ||b - J x|| <= rtol ||b||,
// the outer test of the reference's scipy gmres call.
// ---------------------------------------------------------------------------
// Repeated line Gauss-Seidel sweeps are not a contraction on every system:
// when a restart cycle leaves the residual above half of what it started
// from, the rest of the solve uses the cell block-Jacobi local stage
// (flexible GMRES allows the preconditioner to change between iterations).
static bool stagnated(LinearSolverOptions& lo, double now, double before, int verbose) {
  if (lo.local_stage != 1 || !(now > 0.5 * before)) return false;
  lo.local_stage = 0;
  if (verbose) fprintf(stderr, "[pmx] local stage: line Gauss-Seidel stalled, block-Jacobi for the rest of the solve\n");
  return true;
}

// A solve whose restart cycles have not halved the best residual ten times
// in a row will not converge within the budget: stop it (not converged), so
// the Newton step is cut at once instead of after maxiter cycles.
constexpr int kHopelessCycles = 10;
static bool hopeless(double beta, double& best, int& since_best) {
  if (beta < 0.5 * best) {
    best = beta;
    since_best = 0;
    return false;
  }
  return ++since_best >= kHopelessCycles;
}

// shape codes of the projections: 2, 4, 8 elements per thread with shuffle
// sums; 104, 108: 4 or 8 elements per thread with LDS sums
static int dot2_blocks(size_t n, int e) {
  const size_t tile = (size_t)kRedThreads * (e % 100);
  return (int)((n + tile - 1) / tile);
}

static void multidot2_any(int e, int nblk, hipStream_t s, size_t n, int nv, const double* V, size_t ld,
                          const double* a, const double* w, double* partial) {
  if (e == 104)
    hipLaunchKernelGGL(k_multidot2_lds<4>, dim3(nblk), dim3(kRedThreads), 0, s, n, nv, V, ld, a, w, partial);
  else if (e == 108)
    hipLaunchKernelGGL(k_multidot2_lds<8>, dim3(nblk), dim3(kRedThreads), 0, s, n, nv, V, ld, a, w, partial);
  else if (e == 2)
    hipLaunchKernelGGL(k_multidot2<2>, dim3(nblk), dim3(kRedThreads), 0, s, n, nv, V, ld, a, w, partial, nullptr,
                       nullptr);
  else if (e == 4)
    hipLaunchKernelGGL(k_multidot2<4>, dim3(nblk), dim3(kRedThreads), 0, s, n, nv, V, ld, a, w, partial, nullptr,
                       nullptr);
  else
    hipLaunchKernelGGL(k_multidot2<8>, dim3(nblk), dim3(kRedThreads), 0, s, n, nv, V, ld, a, w, partial, nullptr,
                       nullptr);
  launch_check();
}

__global__ void k_fill_probe(size_t n, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = (double)((i * 2654435761u) % 1000u) / 1000.0 - 0.5;
}

// Elements per thread of the two-sided projections: kDot2Elems, or with
// autotuning the fastest of 2, 4 and 8 with shuffle sums and 4 and 8 with
// LDS sums on this system's length, timed over half the basis on a fixed
// pattern written into the workspace (the solve overwrites it).  Each
// candidate's dots are checked against the default's first; one that
// differs by more than 1e-12 relative is not taken.
static int dot2_shape(CprSolver* C, size_t n, int m, size_t ld, hipStream_t s) {
  if (stream_capturing(s) || m < 4) return kDot2Elems;
  const int nv = m / 2, nd = 2 * nv + 3;
  const TuneKey k{3, 8, octave4((double)n), octave4((double)nv), 0};
  int best = kDot2Elems;
  if (tune_lookup(k, &best)) return best;
  hipLaunchKernelGGL(k_fill_probe, dim3(grid_for(ld * (nv + 2))), dim3(256), 0, s, ld * (nv + 2), C->V.p);
  launch_check();
  auto dots = [&](int e) {
    multidot2_any(e, dot2_blocks(n, e), s, n, nv, C->V.p + 2 * ld, ld, C->V.p, C->V.p + ld, C->partial.p);
    hipLaunchKernelGGL(k_reduce_partials, dim3(nd), dim3(256), 0, s, dot2_blocks(n, e), nd, C->partial.p,
                       C->hdev.p);
    launch_check();
    std::vector<double> h(nd);
    PMX_HIP(hipMemcpyAsync(h.data(), C->hdev.p, sizeof(double) * nd, hipMemcpyDeviceToHost, s));
    PMX_HIP(hipStreamSynchronize(s));
    return h;
  };
  const std::vector<double> ref = dots(kDot2Elems);
  double scale = 0.0;
  for (double v : ref) scale = std::max(scale, std::fabs(v));
  const int cand[5] = {2, 4, 8, 104, 108};
  std::vector<float> t;
  for (int e : cand) {
    const std::vector<double> h = dots(e);
    double err = 0.0;
    for (int i = 0; i < nd; ++i) err = std::max(err, std::fabs(h[i] - ref[i]));
    if (!(err <= 1e-12 * scale)) {
      t.push_back(std::numeric_limits<float>::max());
      continue;
    }
    t.push_back(time_launches(s, 5, [&] {
      multidot2_any(e, dot2_blocks(n, e), s, n, nv, C->V.p + 2 * ld, ld, C->V.p, C->V.p + ld, C->partial.p);
    }));
  }
  best = cand[tune_pick(t, 2)];
  tune_store(k, best);
  return best;
}

int cpr_solve(CprSolver* C, const BlockSystem& S, const double* b, double* x, const LinearSolverOptions& o,
              SolveInfo& info, hipStream_t s) {
  if (C->seg) return cpr_solve_segmented(C, S, b, x, o, info, s);
  auto t0 = std::chrono::steady_clock::now();
  const size_t n = S.n();
  const int m = std::max(1, std::min(o.restart, o.maxiter));
  const size_t ld = (n + 63) / 64 * 64;
  if (C->m_alloc < m || C->n_alloc != n) {
    C->V.alloc(ld * (m + 1));
    C->Z.alloc(ld * m);
    C->w.alloc(ld);
    C->r.alloc(ld);
    // room for the smallest tile of k_multidot2 (kRedThreads * 2 elements)
    C->partial.alloc((size_t)dot2_blocks(n, 2) * (2 * m + 4) + 2 * (2 * m + 4));
    C->hdev.alloc(2 * m + 4);
    C->ydev.alloc(2 * m + 4);
    C->hmat.alloc((size_t)std::max(S.nw, 1) * kMaxMix);
    C->m_alloc = m;
    C->n_alloc = n;
    C->ld = ld;
  }
  double* hscr = C->hmat.p;
  const int nb = grid_for(n);
  const int rb = red_blocks(n);
  if (o.dot_shape > 0) {
    PMX_CHECK(o.dot_shape == 2 || o.dot_shape == 4 || o.dot_shape == 8 || o.dot_shape == 104 ||
                  o.dot_shape == 108,
              "dot_shape is 0 or one of 2, 4, 8, 104, 108");
    C->dot2_elems = o.dot_shape;
    C->dot2_n = 0;   // choose again when the forced shape is lifted
  } else if (C->dot2_elems == 0 || C->dot2_n != n || (!o.autotune && C->dot2_elems != kDot2Elems)) {
    C->dot2_elems = o.autotune ? dot2_shape(C, n, m, ld, s) : kDot2Elems;
    C->dot2_n = n;
  }
  const int e2dot = C->dot2_elems;
  const int rb2 = dot2_blocks(n, e2dot);
  // ||b||
  const double bnorm = std::sqrt(std::max(0.0, device_dot(b, b, n, C->partial.p, s)));
  PMX_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
  info = SolveInfo();
  if (!(bnorm > 0.0)) {
    info.rel_residual = 0.0;
    return std::isfinite(bnorm) ? 0 : -1;
  }
  const double tol = o.rtol * bnorm;
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), hcol(m + 2), y(m);
  double beta = bnorm, beta_cycle = bnorm, best = bnorm;
  int since_best = 0;
  LinearSolverOptions lo = o;   // the local stage may fall back below
  // r = b
  PMX_HIP(hipMemcpyAsync(C->r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  int total = 0;
  int status = 1;
  int cycles = 0;
  // maxiter counts restart cycles, as scipy's gmres does (the reference's
  // CPRSolver calls gmres(restart=50, maxiter=200))
  while (cycles < o.maxiter) {
    ++cycles;
    // V0 = r / beta
    hipLaunchKernelGGL(k_scale_copy, dim3(nb), dim3(256), 0, s, n, 1.0 / beta, C->r.p, C->V.p);
    launch_check();
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    // Arnoldi with delayed classical Gram-Schmidt reorthogonalisation: the
    // new vector w = J z_j is projected once against the basis; slot j+1
    // keeps that "lagged" remainder (scaled by sig), and it is
    // reorthogonalised one iteration later in the same reduction that
    // projects the next w (k_multidot2 + k_dcgs2_update: two passes over
    // the basis per iteration).  Column j of the Hessenberg matrix is
    // completed then; a tentative residual estimate from the lagged norm
    // decides when to complete the last column early (one extra reduction)
    // rather than spend a preconditioner application.  With
    // reorthogonalize = 0 this is classical Gram-Schmidt, once.
    const bool delayed = o.reorthogonalize != 0;
    auto rotate = [&](int col) -> double {
      for (int i = 0; i < col; ++i) {
        const double a = H[(size_t)i * m + col], bq = H[(size_t)(i + 1) * m + col];
        H[(size_t)i * m + col] = cs[i] * a + sn[i] * bq;
        H[(size_t)(i + 1) * m + col] = -sn[i] * a + cs[i] * bq;
      }
      const double a = H[(size_t)col * m + col], bq = H[(size_t)(col + 1) * m + col];
      const double rr = std::hypot(a, bq);
      cs[col] = rr > 0 ? a / rr : 1.0;
      sn[col] = rr > 0 ? bq / rr : 0.0;
      H[(size_t)col * m + col] = rr;
      H[(size_t)(col + 1) * m + col] = 0.0;
      g[col + 1] = -sn[col] * g[col];
      g[col] = cs[col] * g[col];
      return std::fabs(g[col + 1]);
    };
    auto reduce_to_host = [&](int nd, std::vector<double>& hh, int nblk) {
      hipLaunchKernelGGL(k_reduce_partials, dim3(nd), dim3(256), 0, s, nblk, nd, C->partial.p, C->hdev.p);
      launch_check();
      hh.assign(nd, 0.0);
      PMX_HIP(hipMemcpyAsync(hh.data(), C->hdev.p, sizeof(double) * nd, hipMemcpyDeviceToHost, s));
      PMX_HIP(hipStreamSynchronize(s));
    };
    // the lagged norm from Pythagoras; where it cancels, scale by ||w|| (the
    // scale only sets the size of the stored vector)
    auto lag_scale = [](double e2, double ww) {
      if (e2 > 1e-12 * ww) return std::sqrt(e2);
      return ww > 0.0 ? std::sqrt(ww) : 1.0;
    };
    int j = 0, k = 0;
    double resid = beta;
    bool final = true;   // slot j holds an orthonormalised vector
    double sig = 1.0;    // otherwise slot j = sig * (remainder of column j-1)
    std::vector<double> hh, coef;
    for (; j < m; ++j) {
      double* vj = C->V.p + (size_t)j * ld;
      double* zj = C->Z.p + (size_t)j * ld;
      double* w = C->V.p + (size_t)(j + 1) * ld;
      C->prec_t.start(s);
      cpr_apply(C, S, vj, zj, lo, s);
      C->prec_t.stop(s);
      system_matvec(S, zj, w, hscr, s, &C->spmv_t);
      ++total;
      double e2, ww, est;
      if (final) {
        hipLaunchKernelGGL(k_multidot, dim3(rb), dim3(kRedThreads), 0, s, n, j + 1, C->V.p, ld, w, C->partial.p);
        launch_check();
        reduce_to_host(j + 2, hh, rb);
        double hsq = 0.0;
        for (int i = 0; i <= j; ++i) {
          H[(size_t)i * m + j] = hh[i];
          hsq += hh[i] * hh[i];
        }
        ww = hh[j + 1];
        e2 = ww - hsq;
        if (!std::isfinite(ww)) { status = -1; break; }
        est = lag_scale(e2, ww);
        hipLaunchKernelGGL(k_multiaxpy_scale, dim3(nb), dim3(256), 0, s, n, j + 1, C->V.p, ld, C->hdev.p,
                           1.0 / est, w);
        launch_check();
      } else {
        const int nd = 2 * j + 3;
        multidot2_any(e2dot, rb2, s, n, j, C->V.p, ld, vj, w, C->partial.p);
        reduce_to_host(nd, hh, rb2);
        const double* sv = hh.data();
        const double* tv = hh.data() + j;
        const double alpha = hh[2 * j], bp = hh[2 * j + 1];
        ww = hh[2 * j + 2];
        if (!std::isfinite(ww) || !std::isfinite(alpha)) { status = -1; break; }
        double ss = 0.0, st = 0.0, tt = 0.0;
        for (int i = 0; i < j; ++i) {
          ss += sv[i] * sv[i];
          st += sv[i] * tv[i];
          tt += tv[i] * tv[i];
        }
        const double nu = std::sqrt(std::max(alpha - ss, 0.0));
        // complete column j-1: its remainder is slot j / sig = (V s + nu v_j) / sig
        for (int i = 0; i < j; ++i) H[(size_t)i * m + j - 1] += sv[i] / sig;
        H[(size_t)j * m + j - 1] = nu / sig;
        ++info.reorthogonalizations;
        resid = rotate(j - 1);
        if (resid <= tol || !(nu > 0.0)) { k = j; break; }   // this w is not needed
        const double hjj = (bp - st) / nu;
        for (int i = 0; i < j; ++i) H[(size_t)i * m + j] = tv[i];
        H[(size_t)j * m + j] = hjj;
        e2 = ww - tt - hjj * hjj;
        est = lag_scale(e2, ww);
        coef.resize(2 * (size_t)j);
        for (int i = 0; i < j; ++i) {
          coef[i] = sv[i];
          coef[j + i] = tv[i] - sv[i] * hjj / nu;
        }
        PMX_HIP(hipMemcpyAsync(C->ydev.p, coef.data(), sizeof(double) * 2 * j, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_dcgs2_update, dim3(nb), dim3(256), 0, s, n, j, C->V.p, ld, C->ydev.p, 1.0 / nu,
                           hjj / nu, 1.0 / est, vj, w);
        launch_check();
      }
      H[(size_t)(j + 1) * m + j] = est;
      if (!delayed) {
        // classical Gram-Schmidt once: est is taken as the remainder's norm
        resid = rotate(j);
        if (o.verbose > 1) fprintf(stderr, "[pmx] fgmres it %d  rel %.3e\n", total, resid / bnorm);
        k = j + 1;
        if (resid <= tol || !(e2 > 0.0)) break;
        continue;
      }
      sig = 1.0 / est;
      final = false;
      // tentative residual with the lagged norm
      for (int i = 0; i <= j + 1; ++i) hcol[i] = H[(size_t)i * m + j];
      for (int i = 0; i < j; ++i) {
        const double a = hcol[i], bq = hcol[i + 1];
        hcol[i] = cs[i] * a + sn[i] * bq;
        hcol[i + 1] = -sn[i] * a + cs[i] * bq;
      }
      const double rr = std::hypot(hcol[j], hcol[j + 1]);
      const double tent = rr > 0 ? std::fabs(hcol[j + 1] / rr * g[j]) : 0.0;
      if (o.verbose > 1) fprintf(stderr, "[pmx] fgmres it %d  rel %.3e (lagged)\n", total, tent / bnorm);
      if (tent <= tol || j == m - 1) {
        // complete column j now: reorthogonalise slot j+1 by itself
        hipLaunchKernelGGL(k_multidot, dim3(rb), dim3(kRedThreads), 0, s, n, j + 1, C->V.p, ld, w, C->partial.p);
        launch_check();
        reduce_to_host(j + 2, hh, rb);
        double ss = 0.0;
        for (int i = 0; i <= j; ++i) ss += hh[i] * hh[i];
        const double nu = std::sqrt(std::max(hh[j + 1] - ss, 0.0));
        for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] += hh[i] / sig;
        H[(size_t)(j + 1) * m + j] = nu / sig;
        ++info.reorthogonalizations;
        resid = rotate(j);
        k = j + 1;
        if (resid <= tol || j == m - 1 || !(nu > 0.0)) break;
        hipLaunchKernelGGL(k_multiaxpy_scale, dim3(nb), dim3(256), 0, s, n, j + 1, C->V.p, ld, C->hdev.p,
                           1.0 / nu, w);
        launch_check();
        final = true;
        sig = 1.0;
      }
    }
    j = k;
    if (status == -1) break;
    // back-substitution y = R^{-1} g
    for (int i = j - 1; i >= 0; --i) {
      double t = g[i];
      for (int k = i + 1; k < j; ++k) t -= H[(size_t)i * m + k] * y[k];
      y[i] = t / H[(size_t)i * m + i];
    }
    PMX_HIP(hipMemcpyAsync(C->ydev.p, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_multiaxpy_add, dim3(nb), dim3(256), 0, s, n, j, C->Z.p, ld, C->ydev.p, x);
    launch_check();
    // true residual
    system_matvec(S, x, C->r.p, hscr, s);
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, s, n, 1.0, b, -1.0, C->r.p);
    launch_check();
    beta = std::sqrt(std::max(0.0, device_dot(C->r.p, C->r.p, n, C->partial.p, s)));
    info.rel_residual = beta / bnorm;
    if (!std::isfinite(beta)) { status = -1; break; }
    if (beta <= tol) { status = 0; break; }
    ++info.restarts;
    stagnated(lo, beta, beta_cycle, o.verbose);
    beta_cycle = beta;
    if (hopeless(beta, best, since_best)) break;
  }
  info.iterations = total;
  info.levels = (int)C->L.size();
  info.coarse_n = C->coarse_n;
  info.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (o.verbose)
    fprintf(stderr, "[pmx] fgmres (FP%d preconditioner): %d its (%d restarts), rel residual %.2e, %.2f ms, status %d\n",
            C->prec_setup, total, info.restarts, info.rel_residual, info.solve_ms, status);
  return status;
}

}  // namespace pmx

// ---------------------------------------------------------------------------
// Segmented FGMRES: independent right-hand sides solved together.
//
// The unknowns are split into segments (the members of an ensemble, each a
// range of cell unknowns and a range of well unknowns).  The operator and
// the preconditioner are block-diagonal over segments, so one launch of
// each serves all of them, but every segment keeps its own Krylov
// coefficients: its own Gram-Schmidt dots, Hessenberg matrix, Givens
// rotations, convergence test (||b_s - J x_s|| <= rtol ||b_s||) and
// restarts.  A single Krylov space over the union would have to damp every
// segment's outlying eigenvalues with one polynomial, and stalls as the
// segments multiply.  Reductions run over tiles that never straddle two
// segments; a segment's sum is its tiles' partial sums in tile order, so
// the result does not depend on timing.
// ---------------------------------------------------------------------------
namespace pmx {

struct SegmentPlan {
  int nseg = 0, ntile = 0;
  DArr<int> tile_start, tile_len, tptr, tidx, seg_of;
  std::vector<int> h_tptr;
};

__global__ __launch_bounds__(kRedThreads) void k_seg_multidot(int ntile, const int* __restrict__ tstart,
                                                              const int* __restrict__ tlen, int nv,
                                                              const double* __restrict__ V, size_t ld,
                                                              const double* __restrict__ w,
                                                              double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64][kRedGroup];
  const int tile = blockIdx.x;
  if (tile >= ntile) return;
  const int nd = nv + 1;
  const size_t base = (size_t)tstart[tile];
  const int len = tlen[tile];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double wv[kRedElems];
#pragma unroll
  for (int e = 0; e < kRedElems; ++e) {
    const int i = threadIdx.x + e * kRedThreads;
    wv[e] = i < len ? w[base + i] : 0.0;
  }
  for (int d0 = 0; d0 < nd; d0 += kRedGroup) {
    double acc[kRedGroup];
#pragma unroll
    for (int q = 0; q < kRedGroup; ++q) {
      const int d = d0 + q;
      double a = 0.0;
      if (d < nv) {
#pragma unroll
        for (int e = 0; e < kRedElems; ++e) {
          const int i = threadIdx.x + e * kRedThreads;
          if (i < len) a = fma(V[(size_t)d * ld + base + i], wv[e], a);
        }
      } else if (d == nv) {
#pragma unroll
        for (int e = 0; e < kRedElems; ++e) a = fma(wv[e], wv[e], a);
      }
      acc[q] = wave_sum(a, 64);
    }
    if (lane == 0)
#pragma unroll
      for (int q = 0; q < kRedGroup; ++q) sh[wid][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < kRedGroup && d0 + (int)threadIdx.x < nd) {
      double t = 0.0;
#pragma unroll
      for (int v = 0; v < kRedThreads / 64; ++v) t += sh[v][threadIdx.x];
      partial[(size_t)tile * nd + d0 + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

// out[s*nd + d] = sum over segment s's tiles (in order) of partial[tile*nd + d]
__global__ void k_seg_reduce(int nseg, int nd, const int* __restrict__ tptr, const int* __restrict__ tidx,
                             const double* __restrict__ partial, double* __restrict__ out) {
  const int s = blockIdx.x;
  for (int d = threadIdx.x; d < nd; d += blockDim.x) {
    double t = 0.0;
    for (int k = tptr[s]; k < tptr[s + 1]; ++k) t += partial[(size_t)tidx[k] * nd + d];
    out[(size_t)s * nd + d] = t;
  }
}

// w -= sum_d coef[seg(i)*ldc + d] V_d (per tile, so partial[tile] = |w|^2 over the tile)
__global__ __launch_bounds__(kRedThreads) void k_seg_multiaxpy_norm(int ntile, const int* __restrict__ tstart,
                                                                    const int* __restrict__ tlen,
                                                                    const int* __restrict__ seg_of, int nv,
                                                                    const double* __restrict__ V, size_t ld,
                                                                    const double* __restrict__ coef, int ldc,
                                                                    double* __restrict__ w,
                                                                    double* __restrict__ partial) {
  __shared__ double sh[kRedThreads / 64];
  const int tile = blockIdx.x;
  if (tile >= ntile) return;
  const size_t base = (size_t)tstart[tile];
  const int len = tlen[tile];
  const int sg = seg_of[base];
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < kRedElems; ++e) {
    const int i = threadIdx.x + e * kRedThreads;
    if (i < len) {
      double wi = w[base + i];
      for (int d = 0; d < nv; ++d) wi = fma(-coef[(size_t)sg * ldc + d], V[(size_t)d * ld + base + i], wi);
      w[base + i] = wi;
      acc = fma(wi, wi, acc);
    }
  }
  double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[tile] = s;
}

// w = (w - sum_d coef[seg,d] V_d) * scale[seg]
__global__ void k_seg_multiaxpy_scale(size_t n, const int* __restrict__ seg_of, int nv, const double* __restrict__ V,
                                      size_t ld, const double* __restrict__ coef, int ldc,
                                      const double* __restrict__ scale, double* __restrict__ w) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sg = seg_of[i];
  double wi = w[i];
  for (int d = 0; d < nv; ++d) wi = fma(-coef[(size_t)sg * ldc + d], V[(size_t)d * ld + i], wi);
  w[i] = wi * scale[sg];
}

// x += sum_d coef[seg,d] Z_d
__global__ void k_seg_multiaxpy_add(size_t n, const int* __restrict__ seg_of, int nv, const double* __restrict__ Z,
                                    size_t ld, const double* __restrict__ coef, int ldc, double* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sg = seg_of[i];
  double xi = x[i];
  for (int d = 0; d < nv; ++d) xi = fma(coef[(size_t)sg * ldc + d], Z[(size_t)d * ld + i], xi);
  x[i] = xi;
}

// Delayed reorthogonalisation step per segment (k_dcgs2_update with the
// coefficients of the element's segment: s at coef[seg*ldc], c at
// coef[seg*ldc + nv], and scalars sc[seg] = (1/nu, h_jj/nu, 1/est)).
__global__ void k_seg_dcgs2_update(size_t n, const int* __restrict__ seg_of, int nv, const double* __restrict__ V,
                                   size_t ld, const double* __restrict__ coef, int ldc,
                                   const double* __restrict__ sc, double* __restrict__ a, double* __restrict__ w) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sg = seg_of[i];
  const double* cf = coef + (size_t)sg * ldc;
  const double av = a[i];
  double sa = 0.0, sw = 0.0;
  for (int d = 0; d < nv; ++d) {
    const double v = V[(size_t)d * ld + i];
    sa = fma(cf[d], v, sa);
    sw = fma(cf[nv + d], v, sw);
  }
  a[i] = (av - sa) * sc[3 * sg];
  w[i] = (w[i] - sw - sc[3 * sg + 1] * av) * sc[3 * sg + 2];
}

__global__ void k_seg_scale_copy(size_t n, const int* __restrict__ seg_of, const double* __restrict__ scale,
                                 const double* __restrict__ x, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = scale[seg_of[i]] * x[i];
}

struct SegFGMRES {
  SegmentPlan plan;
  DArr<double> partial, dots, coef, scale, scal3;
};

void cpr_set_segments(CprSolver* C, int nseg, const int* cell_ptr, const int* well_ptr, int B, int M, int nc,
                      int nw) {
  PMX_CHECK(nseg >= 1, "need at least one segment");
  PMX_CHECK(cell_ptr[0] == 0 && cell_ptr[nseg] == nc, "segment cell ranges must cover the cells");
  PMX_CHECK(well_ptr[0] == 0 && well_ptr[nseg] == nw, "segment well ranges must cover the wells");
  if (!C->seg) C->seg = std::make_unique<SegFGMRES>();
  SegmentPlan& P = C->seg->plan;
  const size_t n = (size_t)nc * B + (size_t)nw * M;
  std::vector<int> ts, tl, tsg, seg_of(n);
  auto add_range = [&](size_t a, size_t b, int s) {
    for (size_t i = a; i < b; ++i) seg_of[i] = s;
    for (size_t t = a; t < b; t += kRedTile) {
      ts.push_back((int)t);
      tl.push_back((int)std::min<size_t>(kRedTile, b - t));
      tsg.push_back(s);
    }
  };
  for (int s = 0; s < nseg; ++s) {
    PMX_CHECK(cell_ptr[s + 1] >= cell_ptr[s] && well_ptr[s + 1] >= well_ptr[s], "bad segment ranges");
    add_range((size_t)cell_ptr[s] * B, (size_t)cell_ptr[s + 1] * B, s);
    add_range((size_t)nc * B + (size_t)well_ptr[s] * M, (size_t)nc * B + (size_t)well_ptr[s + 1] * M, s);
  }
  const int nt = (int)ts.size();
  std::vector<int> tptr(nseg + 1, 0), tidx(nt);
  for (int t = 0; t < nt; ++t) tptr[tsg[t] + 1]++;
  for (int s = 0; s < nseg; ++s) tptr[s + 1] += tptr[s];
  std::vector<int> fill(tptr.begin(), tptr.end() - 1);
  for (int t = 0; t < nt; ++t) tidx[fill[tsg[t]]++] = t;
  hipStream_t s0 = nullptr;
  P.nseg = nseg;
  P.ntile = nt;
  P.tile_start.upload(ts, s0);
  P.tile_len.upload(tl, s0);
  P.tptr.upload(tptr, s0);
  P.tidx.upload(tidx, s0);
  P.seg_of.upload(seg_of, s0);
  P.h_tptr = tptr;
  PMX_HIP(hipStreamSynchronize(s0));
}

int cpr_solve_segmented(CprSolver* C, const BlockSystem& S, const double* b, double* x,
                        const LinearSolverOptions& o, SolveInfo& info, hipStream_t s) {
  auto t0 = std::chrono::steady_clock::now();
  SegFGMRES& F = *C->seg;
  SegmentPlan& P = F.plan;
  const int ns = P.nseg;
  const size_t n = S.n();
  PMX_CHECK(P.seg_of.n == n, "segments do not match the system");
  const int m = std::max(1, std::min(o.restart, o.maxiter));
  const size_t ld = (n + 63) / 64 * 64;
  if (C->m_alloc < m || C->n_alloc != n) {
    C->V.alloc(ld * (m + 1));
    C->Z.alloc(ld * m);
    C->r.alloc(ld);
    C->hmat.alloc((size_t)std::max(S.nw, 1) * kMaxMix);
    C->m_alloc = m;
    C->n_alloc = n;
    C->ld = ld;
  }
  const int ldc = 2 * m + 2;   // per-segment coefficient stride
  F.partial.alloc((size_t)P.ntile * (2 * m + 4));
  F.dots.alloc((size_t)ns * (2 * m + 4));
  F.coef.alloc((size_t)ns * ldc);
  F.scale.alloc(ns);
  F.scal3.alloc((size_t)ns * 3);
  double* hscr = C->hmat.p;
  const int nb = grid_for(n);
  auto seg_dots = [&](int nv, const double* w, std::vector<double>& out) {
    const int nd = nv + 1;
    hipLaunchKernelGGL(k_seg_multidot, dim3(P.ntile), dim3(kRedThreads), 0, s, P.ntile, P.tile_start.p,
                       P.tile_len.p, nv, C->V.p, ld, w, F.partial.p);
    launch_check();
    hipLaunchKernelGGL(k_seg_reduce, dim3(ns), dim3(64), 0, s, ns, nd, P.tptr.p, P.tidx.p, F.partial.p, F.dots.p);
    launch_check();
    out.resize((size_t)ns * nd);
    PMX_HIP(hipMemcpyAsync(out.data(), F.dots.p, sizeof(double) * ns * nd, hipMemcpyDeviceToHost, s));
    PMX_HIP(hipStreamSynchronize(s));
  };
  // ||b_s|| (V holds nothing yet: nv = 0 gives only w.w)
  std::vector<double> dd;
  seg_dots(0, b, dd);
  std::vector<double> bnorm(ns), tol(ns), beta(ns);
  for (int k = 0; k < ns; ++k) {
    bnorm[k] = std::sqrt(std::max(0.0, dd[k]));
    tol[k] = o.rtol * bnorm[k];
  }
  PMX_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
  info = SolveInfo();
  std::vector<char> fin(ns, 0);   // converged for good
  for (int k = 0; k < ns; ++k) {
    if (!std::isfinite(bnorm[k])) return -1;
    fin[k] = !(bnorm[k] > 0.0);
    beta[k] = bnorm[k];
  }
  PMX_HIP(hipMemcpyAsync(C->r.p, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  std::vector<double> H((size_t)ns * (m + 1) * m), cs((size_t)ns * m), sn((size_t)ns * m), g((size_t)ns * (m + 1)),
      y((size_t)ns * m), hscale(ns), hcoef((size_t)ns * ldc);
  std::vector<int> kdone(ns);
  int total = 0, status = 1, cycles = 0;
  auto all_fin = [&]() {
    for (int k = 0; k < ns; ++k)
      if (!fin[k]) return false;
    return true;
  };
  double worst = 0.0, worst_cycle = 1.0, worst_best = 1.0;
  int since_best = 0;
  int worst_seg = -1;
  LinearSolverOptions lo = o;   // the local stage may fall back below
  while (cycles < o.maxiter && !all_fin()) {
    ++cycles;
    for (int k = 0; k < ns; ++k) hscale[k] = fin[k] ? 0.0 : 1.0 / beta[k];
    PMX_HIP(hipMemcpyAsync(F.scale.p, hscale.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_seg_scale_copy, dim3(nb), dim3(256), 0, s, n, P.seg_of.p, F.scale.p, C->r.p, C->V.p);
    launch_check();
    std::fill(g.begin(), g.end(), 0.0);
    std::vector<char> active(ns);
    for (int k = 0; k < ns; ++k) {
      active[k] = !fin[k];
      g[(size_t)k * (m + 1)] = beta[k];
      kdone[k] = 0;
    }
    // Delayed classical Gram-Schmidt per segment (as cpr_solve): slot j+1
    // keeps the lagged remainder of column j, reorthogonalised in the next
    // iteration's reduction, where column j is completed and each segment's
    // convergence is tested.  A segment that converges stops; its slots are
    // zeroed.  The last column of a cycle is completed by one extra
    // reduction.
    const bool delayed = o.reorthogonalize != 0;
    auto rotate = [&](int k, int col) -> double {
      double* Hk = &H[(size_t)k * (m + 1) * m];
      double* csk = &cs[(size_t)k * m];
      double* snk = &sn[(size_t)k * m];
      double* gk = &g[(size_t)k * (m + 1)];
      for (int i = 0; i < col; ++i) {
        const double a = Hk[(size_t)i * m + col], bq = Hk[(size_t)(i + 1) * m + col];
        Hk[(size_t)i * m + col] = csk[i] * a + snk[i] * bq;
        Hk[(size_t)(i + 1) * m + col] = -snk[i] * a + csk[i] * bq;
      }
      const double a = Hk[(size_t)col * m + col], bq = Hk[(size_t)(col + 1) * m + col];
      const double rr = std::hypot(a, bq);
      csk[col] = rr > 0 ? a / rr : 1.0;
      snk[col] = rr > 0 ? bq / rr : 0.0;
      Hk[(size_t)col * m + col] = rr;
      Hk[(size_t)(col + 1) * m + col] = 0.0;
      gk[col + 1] = -snk[col] * gk[col];
      gk[col] = csk[col] * gk[col];
      return std::fabs(gk[col + 1]);
    };
    auto lag_scale = [](double e2, double ww) {
      if (e2 > 1e-12 * ww) return std::sqrt(e2);
      return ww > 0.0 ? std::sqrt(ww) : 1.0;
    };
    auto Hs = [&](int k, int i, int c) -> double& { return H[(size_t)k * (m + 1) * m + (size_t)i * m + c]; };
    std::vector<double> sig(ns, 1.0), sc3((size_t)ns * 3);
    bool final = true;
    int j = 0;
    for (; j < m; ++j) {
      double* vj = C->V.p + (size_t)j * ld;
      double* zj = C->Z.p + (size_t)j * ld;
      double* w = C->V.p + (size_t)(j + 1) * ld;
      C->prec_t.start(s);
      cpr_apply(C, S, vj, zj, lo, s);
      C->prec_t.stop(s);
      system_matvec(S, zj, w, hscr, s, &C->spmv_t);
      ++total;
      bool any_active = false;
      std::fill(hcoef.begin(), hcoef.end(), 0.0);
      if (final || !delayed) {
        const int nd = j + 2;
        std::vector<double> hh;
        seg_dots(j + 1, w, hh);
        for (int k = 0; k < ns; ++k) {
          hscale[k] = 0.0;
          if (!active[k]) continue;
          double hsq = 0.0;
          for (int i = 0; i <= j; ++i) {
            const double h = hh[(size_t)k * nd + i];
            Hs(k, i, j) = h;
            hcoef[(size_t)k * ldc + i] = h;
            hsq += h * h;
          }
          const double ww = hh[(size_t)k * nd + j + 1];
          if (!std::isfinite(ww)) return -1;
          const double e2 = ww - hsq, est = lag_scale(e2, ww);
          Hs(k, j + 1, j) = est;
          hscale[k] = 1.0 / est;
          if (!delayed) {
            // classical Gram-Schmidt once: est is the remainder's norm
            kdone[k] = j + 1;
            if (rotate(k, j) <= tol[k] || !(e2 > 0.0)) {
              active[k] = 0;
              hscale[k] = 0.0;
              continue;
            }
          } else {
            sig[k] = 1.0 / est;
          }
          any_active = true;
        }
        if (!any_active) {
          ++j;
          break;
        }
        PMX_HIP(hipMemcpyAsync(F.coef.p, hcoef.data(), sizeof(double) * ns * ldc, hipMemcpyHostToDevice, s));
        PMX_HIP(hipMemcpyAsync(F.scale.p, hscale.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_seg_multiaxpy_scale, dim3(nb), dim3(256), 0, s, n, P.seg_of.p, j + 1, C->V.p, ld,
                           F.coef.p, ldc, F.scale.p, w);
        launch_check();
        if (delayed) final = false;
      } else {
        const int nd = 2 * j + 3;
        hipLaunchKernelGGL(k_multidot2<kRedElems>, dim3(P.ntile), dim3(kRedThreads), 0, s, n, j, C->V.p, ld, vj, w,
                           F.partial.p, P.tile_start.p, P.tile_len.p);
        launch_check();
        hipLaunchKernelGGL(k_seg_reduce, dim3(ns), dim3(64), 0, s, ns, nd, P.tptr.p, P.tidx.p, F.partial.p,
                           F.dots.p);
        launch_check();
        std::vector<double> hh((size_t)ns * nd);
        PMX_HIP(hipMemcpyAsync(hh.data(), F.dots.p, sizeof(double) * ns * nd, hipMemcpyDeviceToHost, s));
        PMX_HIP(hipStreamSynchronize(s));
        ++info.reorthogonalizations;
        for (int k = 0; k < ns; ++k) {
          double* s3 = &sc3[(size_t)k * 3];
          s3[0] = s3[1] = s3[2] = 0.0;
          if (!active[k]) continue;
          const double* sv = &hh[(size_t)k * nd];
          const double* tv = sv + j;
          const double alpha = sv[2 * j], bp = sv[2 * j + 1], ww = sv[2 * j + 2];
          if (!std::isfinite(ww) || !std::isfinite(alpha)) return -1;
          double ss = 0.0, st = 0.0, tt = 0.0;
          for (int i = 0; i < j; ++i) {
            ss += sv[i] * sv[i];
            st += sv[i] * tv[i];
            tt += tv[i] * tv[i];
          }
          const double nu = std::sqrt(std::max(alpha - ss, 0.0));
          for (int i = 0; i < j; ++i) Hs(k, i, j - 1) += sv[i] / sig[k];
          Hs(k, j, j - 1) = nu / sig[k];
          kdone[k] = j;
          if (rotate(k, j - 1) <= tol[k] || !(nu > 0.0)) {
            active[k] = 0;   // converged on column j-1; its slots are zeroed below
            continue;
          }
          const double hjj = (bp - st) / nu;
          for (int i = 0; i < j; ++i) Hs(k, i, j) = tv[i];
          Hs(k, j, j) = hjj;
          const double est = lag_scale(ww - tt - hjj * hjj, ww);
          Hs(k, j + 1, j) = est;
          sig[k] = 1.0 / est;
          for (int i = 0; i < j; ++i) {
            hcoef[(size_t)k * ldc + i] = sv[i];
            hcoef[(size_t)k * ldc + j + i] = tv[i] - sv[i] * hjj / nu;
          }
          s3[0] = 1.0 / nu;
          s3[1] = hjj / nu;
          s3[2] = 1.0 / est;
          any_active = true;
        }
        if (!any_active) break;
        PMX_HIP(hipMemcpyAsync(F.coef.p, hcoef.data(), sizeof(double) * ns * ldc, hipMemcpyHostToDevice, s));
        PMX_HIP(hipMemcpyAsync(F.scal3.p, sc3.data(), sizeof(double) * ns * 3, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_seg_dcgs2_update, dim3(nb), dim3(256), 0, s, n, P.seg_of.p, j, C->V.p, ld, F.coef.p,
                           ldc, F.scal3.p, vj, w);
        launch_check();
      }
      if (delayed && j == m - 1) {
        // complete the cycle's last column for the segments still active
        std::vector<double> hh;
        seg_dots(j + 1, w, hh);
        const int nd = j + 2;
        for (int k = 0; k < ns; ++k) {
          if (!active[k]) continue;
          double ss = 0.0;
          for (int i = 0; i <= j; ++i) ss += hh[(size_t)k * nd + i] * hh[(size_t)k * nd + i];
          const double nu = std::sqrt(std::max(hh[(size_t)k * nd + j + 1] - ss, 0.0));
          for (int i = 0; i <= j; ++i) Hs(k, i, j) += hh[(size_t)k * nd + i] / sig[k];
          Hs(k, j + 1, j) = nu / sig[k];
          rotate(k, j);
          kdone[k] = j + 1;
        }
        ++info.reorthogonalizations;
        ++j;
        break;
      }
    }
    for (int k = 0; k < ns; ++k) j = std::max(j, kdone[k]);
    // y_s = R_s^{-1} g_s over each segment's own iterations
    std::fill(hcoef.begin(), hcoef.end(), 0.0);
    for (int k = 0; k < ns; ++k) {
      const int jk = kdone[k];
      const double* Hk = &H[(size_t)k * (m + 1) * m];
      for (int i = jk - 1; i >= 0; --i) {
        double t = g[(size_t)k * (m + 1) + i];
        for (int q = i + 1; q < jk; ++q) t -= Hk[(size_t)i * m + q] * hcoef[(size_t)k * ldc + q];
        hcoef[(size_t)k * ldc + i] = t / Hk[(size_t)i * m + i];
      }
    }
    PMX_HIP(hipMemcpyAsync(F.coef.p, hcoef.data(), sizeof(double) * ns * ldc, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_seg_multiaxpy_add, dim3(nb), dim3(256), 0, s, n, P.seg_of.p, j, C->Z.p, ld, F.coef.p,
                       ldc, x);
    launch_check();
    // true residuals
    system_matvec(S, x, C->r.p, hscr, s);
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(256), 0, s, n, 1.0, b, -1.0, C->r.p);
    launch_check();
    std::vector<double> rr;
    seg_dots(0, C->r.p, rr);
    worst = 0.0;
    for (int k = 0; k < ns; ++k) {
      beta[k] = std::sqrt(std::max(0.0, rr[k]));
      if (!std::isfinite(beta[k])) return -1;
      if (beta[k] <= tol[k]) fin[k] = 1;
      if (bnorm[k] > 0.0 && beta[k] / bnorm[k] > worst) {
        worst = beta[k] / bnorm[k];
        worst_seg = k;
      }
    }
    if (all_fin()) {
      status = 0;
      break;
    }
    ++info.restarts;
    stagnated(lo, worst, worst_cycle, o.verbose);
    worst_cycle = worst;
    if (hopeless(worst, worst_best, since_best)) break;
  }
  info.iterations = total;
  info.rel_residual = worst;
  info.levels = (int)C->L.size();
  info.coarse_n = C->coarse_n;
  info.solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (o.verbose)
    fprintf(stderr, "[pmx] segmented fgmres (%d segments, FP%d preconditioner): %d its (%d restarts), worst rel "
                    "residual %.2e (segment %d, |b| %.3e), %.2f ms, status %d\n",
            ns, C->prec_setup, total, info.restarts, worst, worst_seg, worst_seg >= 0 ? bnorm[worst_seg] : 0.0,
            info.solve_ms, status);
  return status;
}

}  // namespace pmx

namespace pmx {
CprSolver::~CprSolver() {
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (cap_s) (void)hipStreamDestroy(cap_s);
  if (io) (void)hipFree(io);
}
}  // namespace pmx
