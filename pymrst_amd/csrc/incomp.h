// Device incompressible TPFA pressure solve and computeTrans (incomp.hip).
#pragma once
#include "linsys.h"

namespace pmx {

struct IncompBC {
  std::vector<int> face, type;   // type 0: pressure (Dirichlet), 1: flux (Neumann)
  std::vector<double> value;
  std::vector<int> src_cell;
  std::vector<double> src_rate;
  std::vector<int> jump_face;      // periodic boundary: pressure jump added across these faces
  std::vector<double> jump_value;
};

struct IncompWells {
  std::vector<int> perf_ptr, cells, type;   // type 0: bhp, 1: rate
  std::vector<double> WI, val, dp;
};

class IncompDev {
 public:
  explicit IncompDev(hipStream_t s);
  ~IncompDev();
  // face_pos[nc+1] / hf_face[nhf]: the faces of each cell (MRST cells.facePos
  // and cells.faces); face_nb[2*nf] neighbours (-1 outside); centroids.
  void set_grid(int nc, int nf, const int* face_pos, const int* hf_face, const int* face_nb, const double* cc,
                const double* fc);
  void compute_trans(const double* fn, int nk, const double* perm, double* T);
  int solve(const double* T, const double* mob, const double* omega, const double* g, const IncompBC& bc,
            const IncompWells& W, const LinearSolverOptions& o, double* p, double* flux, double* fpress,
            double* well_p, double* well_flux, SolveInfo& info);
  // the stages of solve: assembly, the device solve into x (nc + nw), and
  // the reconstruction from any solution x; export_system copies the
  // assembled system to the host
  void assemble(const double* T, const double* mob, const double* omega, const double* g, const IncompBC& bc,
                const IncompWells& W);
  int solve_assembled(const LinearSolverOptions& o, double* x, SolveInfo& info);
  void reconstruct(const double* x, double* p, double* flux, double* fpress, double* well_p, double* well_flux);
  void export_system(int* rowptr, int* col, double* val, double* Bp, double* Cp, double* D, double* rhs);

  hipStream_t stream;
  int nc = 0, nf = 0, nhf = 0;
  BlockSystem S;
  HostPattern host_pat;
  DArr<int> face_pos, hf_face, hf_cell, face_nb, face_hf;
  DArr<double> cc, fc, thf, ghf, ft, fg, fg_all, dval, hh, gg, r, xw, wi_d, radd_d, w1;
  DArr<unsigned char> dF;
  DArr<int> jump_face;
  DArr<double> jump_val;
  CprSolver* cpr = nullptr;
  // the last assembly's wells and their host-side coefficients
  bool assembled = false;
  IncompWells wells;
  std::vector<double> wi_h, Bp_h, Cp_h, D_h;

 private:
  double* x_work();
};

}  // namespace pmx
