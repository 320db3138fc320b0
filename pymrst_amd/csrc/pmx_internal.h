// Internal declarations shared by the pmx HIP sources (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace pmx {

#define PMX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess)                                                          \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PMX_CHECK(cond, msg)                                     \
  do {                                                           \
    if (!(cond)) throw std::invalid_argument(std::string(msg));  \
  } while (0)

// Owning device array (hipMalloc / hipFree).
template <class T>
struct DArr {
  T* p = nullptr;
  size_t n = 0;
  DArr() = default;
  DArr(const DArr&) = delete;
  DArr& operator=(const DArr&) = delete;
  ~DArr() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t n_) {
    if (n_ == n && p) return;
    release();
    n = n_;
    if (n) PMX_HIP(hipMalloc(&p, n * sizeof(T)));
  }
  void upload(const T* h, size_t n_, hipStream_t s) {
    alloc(n_);
    if (n) PMX_HIP(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void zero(hipStream_t s);
};

// ---------------------------------------------------------------------------
// Host-side pattern building (topology.cpp)
// ---------------------------------------------------------------------------
struct HostPattern {
  int n = 0;                 // block rows
  std::vector<int> rowptr;   // n+1
  std::vector<int> col;      // nnz, sorted per row, diagonal included
  std::vector<int> diag;     // n: position of the diagonal entry
};

// Block pattern of a TPFA operator on n cells with connections N (pairs).
// Also returns, per entry, the connection index*2 + (row is N1) or -1 on the
// diagonal.
void tpfa_pattern(int n, int nconn, const int* N, HostPattern& pat, std::vector<int>& entry_conn);

// One level of the aggregation hierarchy built on the host.
struct HostLevel {
  HostPattern pat;           // this level's pattern
  std::vector<double> w;     // off-diagonal strength weights (per entry, 0 on diag)
  std::vector<int> agg;      // fine -> coarse aggregate (n)
  int ncoarse = 0;
  std::vector<int> agg_ptr, agg_mem;   // coarse -> members
  std::vector<int> gal_ptr, gal_src;   // coarse entry -> fine entries
};

struct AggregationParams {
  double theta = 0.25;   // strength threshold relative to the strongest connection
  int max_levels = 12;
  int coarse_size = 256; // stop when at most this many unknowns
  int min_coarse_reduction_pct = 85;  // stop when a level keeps more than this % of rows
  int pairwise_passes = 0;            // 0: greedy strength aggregation; >0: pairwise matching passes
};

// Build the hierarchy from a fine pattern and symmetric off-diagonal weights.
// levels[0] is the fine level; levels.back() is the coarsest (agg empty).
void build_aggregation(const HostPattern& fine, const std::vector<double>& w,
                       const AggregationParams& prm, std::vector<HostLevel>& levels);

// Lines of strongly coupled cells, stored position-major (entry k*nlines + l
// is the k-th cell of line l, -1 past its end); lo/up are the BCSR entries of
// the blocks (c_k, c_{k-1}) and (c_k, c_{k+1}).  Lines are coloured so that
// no two lines of one colour share a matrix entry; colour c owns lines
// [color_ptr[c], color_ptr[c+1]), each range sorted by descending length and
// padded with empty lines to a multiple of LINE_GROUP (one wave per group).
// maxlen is padded to a multiple of LINE_PAD (tables hold maxlen rows).
constexpr int LINE_PAD = 16;
constexpr int LINE_GROUP = 64;
// The couplings of each line cell to cells off its line are listed in noff
// position-major planes (plane h, entry h*maxlen*nlines + k*nlines + l):
// the BCSR entry, the neighbour's table position and the neighbour's colour
// (-1 / 255 where a cell has fewer).  pos[c] is the table position of cell c.
struct HostLines {
  int nlines = 0, maxlen = 0, longest = 0, ncolor = 0, noff = 0;
  std::vector<int> len, cell, lo, up, color_ptr, pos;
  std::vector<unsigned char> cell_color;   // colour of each cell's line
  std::vector<int> off_q, off_nbr;
  std::vector<unsigned char> off_col;
};
void build_lines(const HostPattern& P, const std::vector<double>& w, HostLines& L);
void lines_from_list(const HostPattern& P, int nlines, const int* ptr, const int* cells, HostLines& L);

}  // namespace pmx
