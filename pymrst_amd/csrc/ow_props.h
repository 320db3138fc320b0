// Cell properties and the per-connection phase flux of the oil-water model,
// shared by the assembly (ow.hip) and its adjoint (adjoint.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ow.h"

namespace pmx {

struct CellProps {
  double bw, dbw, bo, dbo;   // shrinkage factors and d/dp
  double mw, dmw, mo, dmo;   // mobilities and d/dsW
};

__device__ __forceinline__ void shrink(double c, double p, double pref, double& b, double& db) {
  if (c == 0.0) {
    b = 1.0;
    db = 0.0;
  } else {
    b = exp(c * (p - pref));
    db = c * b;
  }
}

__device__ __forceinline__ void corey(double s, double n, double& v, double& dv) {
  if (n == 2.0) {
    v = s * s;
    dv = 2.0 * s;
  } else if (n == 1.0) {
    v = s;
    dv = 1.0;
  } else {
    v = pow(s, n);
    dv = n * pow(s, n - 1.0);
  }
}

__device__ __forceinline__ CellProps cell_props(const OWFluid& f, double p, double s) {
  CellProps c;
  shrink(f.c_w, p, f.p_ref, c.bw, c.dbw);
  shrink(f.c_o, p, f.p_ref, c.bo, c.dbo);
  double kw, dkw, ko, dko;
  corey(s, f.n_w, kw, dkw);
  corey(1.0 - s, f.n_o, ko, dko);
  c.mw = kw / f.mu_w;
  c.dmw = dkw / f.mu_w;
  c.mo = ko / f.mu_o;
  c.dmo = -dko / f.mu_o;
  return c;
}

template <class T>
__device__ __forceinline__ T lane_sum(T v, int width) {
  for (int off = width / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, width);
  return v;
}

// One phase's flux across the connection of block k, seen from row r.
// pot = (p_n - p_r) - avg(rho) G with G = g (z_n - z_r); contribution to
// R(r) is C = b_u (-T mob_u pot).  At pot == 0 the upwind cell is N0.
__device__ __forceinline__ void phase_flux(double T, double G, bool isN0, double pr, double pn, double rhoS,
                                           double br, double dbr, double mr, double dmr, double bn, double dbn,
                                           double mn, double dmn, double& C, double& dCpr, double& dCsr,
                                           double& dCpn, double& dCsn) {
  const double rr = rhoS * br, rn = rhoS * bn;
  const double pot = (pn - pr) - (0.5 * rr + 0.5 * rn) * G;
  const double dpot_r = -1.0 - 0.5 * rhoS * dbr * G;
  const double dpot_n = 1.0 - 0.5 * rhoS * dbn * G;
  const bool up_r = isN0 ? (pot <= 0.0) : (pot < 0.0);
  const double bu = up_r ? br : bn;
  const double mu = up_r ? mr : mn;
  const double tm = -T * mu;
  const double v = tm * pot;
  C = bu * v;
  dCpr = bu * tm * dpot_r + (up_r ? dbr * v : 0.0);
  dCpn = bu * tm * dpot_n + (up_r ? 0.0 : dbn * v);
  dCsr = up_r ? bu * (-T * dmr) * pot : 0.0;
  dCsn = up_r ? 0.0 : bu * (-T * dmn) * pot;
}

}  // namespace pmx
