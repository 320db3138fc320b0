// Corner-point grid processing.
//
// Topology (host): the faces and cells MRST's processGRDECL makes
// (pymrst/gridprocessing/process_grdecl.py:352-580) from pillars (COORD),
// corner depths (ZCORN) and ACTNUM:
//   * lateral faces on every pillar pair: where the two cell stacks along
//     the pair match corner for corner, one face per layer; where they do
//     not (a fault), the stacks of active cells and the gaps between them
//     are intersected pairwise, each overlap a polygon of its own;
//   * vertical faces in every column, consecutive identical corner layers
//     merged, so a layer of zero thickness (a pinch) joins the cells above
//     and below it and vanishes;
//   * a cell is kept when it is active and has a face; a face when it has a
//     kept cell on one side.
// Every face's corners are ordered so that its area vector points from its
// first cell to its second.
//
// Geometry (device): MRST's computeGeometry
// (pymrst/gridprocessing/compute_geometry.py:211-325): each face is a fan
// of triangles around the mean of its corners (area = sum of the triangle
// areas, normal = sum of their area vectors, centroid = area-weighted), each
// cell a set of tetrahedra from the mean of its face centroids to those
// triangles.
#include "cpgrid.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>

namespace pmx {
namespace {

using Vec3 = std::array<double, 3>;

// A face corner: a point on a pillar at a depth, or a point where two cell
// edges of a fault cross (pillar < 0, coordinates given).
struct Corner {
  int pillar;
  double d;
  Vec3 x;
};

// One interval of a cell stack along a pillar pair: the side face of a cell
// (cell >= 0, its Cartesian index) or a gap between cell faces (cell = -1),
// by its top and bottom depth on the first and the second pillar.
struct Interval {
  int cell;
  double t1, b1, t2, b2;
};

struct RawFace {
  std::vector<Corner> corners;
  int c0, c1;    // Cartesian cells, -1 outside
  int tag;
};

// Line labels of a clipped overlap polygon's edges (where a corner found on
// an edge lies): the first or second pillar, or the top or bottom edge of
// side A's or side B's interval.
enum Line { kP0, kP1, kATop, kABot, kBTop, kBBot };

struct PolyVert {
  double s, d;   // position along the pillar pair (0: first pillar) and depth
  int on;        // the line this corner lies on
  int edge;      // the line of the edge from this corner to the next
};

class Processor {
 public:
  Processor(int nx, int ny, int nz, const double* coord, const double* zcorn, const int* act)
      : nx_(nx), ny_(ny), nz_(nz), coord_(coord), zc_(zcorn), act_(act) {
    // depth increases downwards through active cells (else the deck is in
    // elevations: flip for processing, as processGRDECL does)
    bool down = false;
    for (int k2 = 0; k2 + 1 < 2 * nz && !down; ++k2)
      for (int j2 = 0; j2 < 2 * ny && !down; ++j2)
        for (int i2 = 0; i2 < 2 * nx && !down; ++i2) {
          if (!active(i2 / 2, j2 / 2, k2 / 2) || !active(i2 / 2, j2 / 2, (k2 + 1) / 2)) continue;
          if (raw_z(i2, j2, k2 + 1) > raw_z(i2, j2, k2)) down = true;
        }
    sign_ = down ? 1.0 : -1.0;
    dmin_ = INFINITY;
    dmax_ = -INFINITY;
    for (size_t q = 0; q < (size_t)8 * nx * ny * nz; ++q) {
      dmin_ = std::min(dmin_, sign_ * zc_[q]);
      dmax_ = std::max(dmax_, sign_ * zc_[q]);
    }
  }

  void run(CpGrid& g);

 private:
  int nx_, ny_, nz_;
  const double* coord_;
  const double* zc_;
  const int* act_;
  double sign_ = 1.0, dmin_ = 0.0, dmax_ = 0.0;
  std::vector<RawFace> faces_;

  double raw_z(int i2, int j2, int k2) const { return zc_[i2 + (size_t)2 * nx_ * (j2 + (size_t)2 * ny_ * k2)]; }
  double depth(int i2, int j2, int k2) const { return sign_ * raw_z(i2, j2, k2); }
  bool active(int i, int j, int k) const { return act_ == nullptr || act_[i + nx_ * (j + (size_t)ny_ * k)] != 0; }
  int cart(int i, int j, int k) const { return i + nx_ * (j + ny_ * k); }
  int pillar(int i, int j) const { return i + (nx_ + 1) * j; }

  // the point of a pillar at a depth
  Vec3 point(int p, double d) const {
    const double* L = coord_ + 6 * (size_t)p;
    const double z = sign_ * d, dz = L[5] - L[2];
    const double t = std::fabs(dz) < 1e-300 ? 0.0 : (z - L[2]) / dz;
    return {L[0] + t * (L[3] - L[0]), L[1] + t * (L[4] - L[1]), z};
  }
  Corner pc(int p, double d) const { return Corner{p, d, point(p, d)}; }

  void lateral(bool jdir);
  void overlap(const Interval& a, const Interval& b, int p1, int p2, bool jdir);
  void vertical();
};

// Side faces on the pillar pairs of one direction.  i-faces: plane i of
// pillars, pair (i, j)-(i, j+1), cells (i-1, j) on side A and (i, j) on
// side B.  j-faces: plane j, pair (i, j)-(i+1, j), cells (i, j-1) and (i, j).
void Processor::lateral(bool jdir) {
  const int nplane = jdir ? ny_ : nx_;
  const int nrow = jdir ? nx_ : ny_;
  for (int q = 0; q <= nplane; ++q)
    for (int r = 0; r < nrow; ++r) {
      const int p1 = jdir ? pillar(r, q) : pillar(q, r);
      const int p2 = jdir ? pillar(r + 1, q) : pillar(q, r + 1);
      // the depths of layer k's face on side A (s = 0) or B (s = 1)
      auto face = [&](int s, int k, double& t1, double& b1, double& t2, double& b2) {
        const int c = q - 1 + s;                   // cell index along the plane normal
        const int a2 = 2 * c + (s == 0 ? 1 : 0);   // corner index along the normal
        const int r1 = 2 * r, r2 = 2 * r + 1;      // corner indices along the pair
        if (jdir) {
          t1 = depth(r1, a2, 2 * k);
          b1 = depth(r1, a2, 2 * k + 1);
          t2 = depth(r2, a2, 2 * k);
          b2 = depth(r2, a2, 2 * k + 1);
        } else {
          t1 = depth(a2, r1, 2 * k);
          b1 = depth(a2, r1, 2 * k + 1);
          t2 = depth(a2, r2, 2 * k);
          b2 = depth(a2, r2, 2 * k + 1);
        }
      };
      auto cell = [&](int s, int k) -> int {
        const int c = q - 1 + s;
        if (c < 0 || c >= nplane) return -1;
        const int i = jdir ? r : c, j = jdir ? c : r;
        return active(i, j, k) ? cart(i, j, k) : -1;
      };
      const bool hasA = q > 0, hasB = q < nplane;
      bool regular = true;
      if (hasA && hasB)
        for (int k = 0; k < nz_ && regular; ++k) {
          double a[4], b[4];
          face(0, k, a[0], a[1], a[2], a[3]);
          face(1, k, b[0], b[1], b[2], b[3]);
          regular = a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && a[3] == b[3];
        }
      if (regular) {
        for (int k = 0; k < nz_; ++k) {
          const int cA = hasA ? cell(0, k) : -1, cB = hasB ? cell(1, k) : -1;
          if (cA < 0 && cB < 0) continue;
          double t1, b1, t2, b2;
          face(hasA ? 0 : 1, k, t1, b1, t2, b2);
          if (t1 == b1 && t2 == b2) continue;    // zero area: a pinched layer
          RawFace f{{pc(p1, t1), pc(p2, t2), pc(p2, b2), pc(p1, b1)}, cA, cB, jdir ? 2 : 1};
          faces_.push_back(std::move(f));
        }
        continue;
      }
      // a fault: the active cells' faces and the gaps between them on
      // either side, intersected pairwise
      std::vector<Interval> A, B;
      for (int s = 0; s < 2; ++s) {
        std::vector<Interval>& L = s == 0 ? A : B;
        double pb1 = dmin_ - 1.0, pb2 = dmin_ - 1.0;
        for (int k = 0; k < nz_; ++k) {
          const int c = cell(s, k);
          if (c < 0) continue;
          double t1, b1, t2, b2;
          face(s, k, t1, b1, t2, b2);
          L.push_back(Interval{-1, pb1, t1, pb2, t2});
          L.push_back(Interval{c, t1, b1, t2, b2});
          pb1 = b1;
          pb2 = b2;
        }
        L.push_back(Interval{-1, pb1, dmax_ + 1.0, pb2, dmax_ + 1.0});
      }
      for (const Interval& a : A)
        for (const Interval& b : B) {
          if (a.cell < 0 && b.cell < 0) continue;
          // disjoint depth ranges cannot overlap (the clip decides the rest)
          if (std::max(a.b1, a.b2) <= std::min(b.t1, b.t2) || std::max(b.b1, b.b2) <= std::min(a.t1, a.t2))
            continue;
          overlap(a, b, p1, p2, jdir);
        }
    }
}

// The overlap of two intervals on one pillar pair: side A's quadrilateral in
// the (s, depth) plane clipped by side B's top (depth >= top_B(s)) and bottom
// (depth <= bottom_B(s)) lines.  Corners on a pillar are pillar points; a
// corner where the clip lines cross an interval edge lies on that edge,
// placed in space on the straight edge between its two pillar end points
// (side A's edge where it lies on one of A's, as processGRDECL places it).
void Processor::overlap(const Interval& a, const Interval& b, int p1, int p2, bool jdir) {
  std::vector<PolyVert> poly = {{0.0, a.t1, kP0, kATop}, {1.0, a.t2, kP1, kP1}, {1.0, a.b2, kP1, kABot},
                                {0.0, a.b1, kP0, kP0}};
  auto clip = [&](int line, bool keep_below) {
    // keep_below: keep depth >= line(s) (below side B's top); else depth <= line(s)
    const double l1 = line == kBTop ? b.t1 : b.b1, l2 = line == kBTop ? b.t2 : b.b2;
    auto val = [&](const PolyVert& v) {
      const double ls = l1 + v.s * (l2 - l1);
      return keep_below ? v.d - ls : ls - v.d;
    };
    std::vector<PolyVert> out;
    const size_t n = poly.size();
    for (size_t i = 0; i < n; ++i) {
      const PolyVert& u = poly[i];
      const PolyVert& v = poly[(i + 1) % n];
      const double fu = val(u), fv = val(v);
      const bool iu = fu >= 0.0, iv = fv >= 0.0;
      if (iu) out.push_back(u);
      if (iu != iv) {
        const double t = fu / (fu - fv);
        PolyVert x{u.s + t * (v.s - u.s), u.d + t * (v.d - u.d), u.edge, u.edge};
        if (u.edge == kP0) x.s = 0.0;
        if (u.edge == kP1) x.s = 1.0;
        if (iu) {
          x.edge = line;              // leaving: the next edge runs along the clip line
          out.push_back(x);
        } else {
          out.push_back(x);           // entering: the next edge continues along u's edge
        }
      }
    }
    // an entered corner's edge: the edge label of the vertex before it in
    // `out` is the clip line when that vertex was a leaving intersection
    poly.swap(out);
  };
  clip(kBTop, true);
  if (poly.size() < 3) return;
  clip(kBBot, false);
  if (poly.size() < 3) return;
  // drop repeated corners, then measure the area in the (s, depth) plane
  std::vector<PolyVert> q;
  for (size_t i = 0; i < poly.size(); ++i) {
    const PolyVert& v = poly[i];
    if (!q.empty() && q.back().s == v.s && q.back().d == v.d) continue;
    q.push_back(v);
  }
  while (q.size() > 1 && q.front().s == q.back().s && q.front().d == q.back().d) q.pop_back();
  if (q.size() < 3) return;
  double area = 0.0;
  for (size_t i = 0; i < q.size(); ++i) {
    const PolyVert& u = q[i];
    const PolyVert& v = q[(i + 1) % q.size()];
    area += u.s * v.d - v.s * u.d;
  }
  const double scale = std::max({std::fabs(a.b1 - a.t1), std::fabs(a.b2 - a.t2), 1e-300});
  if (!(std::fabs(area) > 1e-12 * scale)) return;
  RawFace f;
  f.c0 = a.cell;
  f.c1 = b.cell;
  f.tag = (jdir ? 2 : 1) + 3;
  for (const PolyVert& v : q) {
    if (v.s == 0.0 || v.on == kP0) {
      f.corners.push_back(pc(p1, v.d));
    } else if (v.s == 1.0 || v.on == kP1) {
      f.corners.push_back(pc(p2, v.d));
    } else {
      // on an interval edge: A's edges where the corner is on one of them
      double e1, e2;
      switch (v.on) {
        case kATop: e1 = a.t1; e2 = a.t2; break;
        case kABot: e1 = a.b1; e2 = a.b2; break;
        case kBTop: e1 = b.t1; e2 = b.t2; break;
        default: e1 = b.b1; e2 = b.b2; break;
      }
      const Vec3 x1 = point(p1, e1), x2 = point(p2, e2);
      Vec3 x;
      for (int d = 0; d < 3; ++d) x[d] = x1[d] + v.s * (x2[d] - x1[d]);
      f.corners.push_back(Corner{-1, v.d, x});
    }
  }
  faces_.push_back(std::move(f));
}

// Faces between layers in every column: consecutive identical corner layers
// form one face (a pinched layer disappears between its neighbours).
void Processor::vertical() {
  for (int j = 0; j < ny_; ++j)
    for (int i = 0; i < nx_; ++i) {
      const int pp[4] = {pillar(i, j), pillar(i + 1, j), pillar(i + 1, j + 1), pillar(i, j + 1)};
      auto layer = [&](int L, double (&d)[4]) {
        d[0] = depth(2 * i, 2 * j, L);
        d[1] = depth(2 * i + 1, 2 * j, L);
        d[2] = depth(2 * i + 1, 2 * j + 1, L);
        d[3] = depth(2 * i, 2 * j + 1, L);
      };
      int L = 0;
      const int nl = 2 * nz_;
      while (L < nl) {
        double d0[4];
        layer(L, d0);
        int e = L + 1;
        for (; e < nl; ++e) {
          double d1[4];
          layer(e, d1);
          if (!(d1[0] == d0[0] && d1[1] == d0[1] && d1[2] == d0[2] && d1[3] == d0[3])) break;
        }
        // the run [L, e): the cell above ends at an odd layer, the cell
        // below starts at an even one
        const int above = (L % 2 == 1) ? (L - 1) / 2 : -1;
        const int below = (e % 2 == 1) ? (e - 1) / 2 : -1;
        const int cA = above >= 0 && active(i, j, above) ? cart(i, j, above) : -1;
        const int cB = below >= 0 && active(i, j, below) ? cart(i, j, below) : -1;
        if (cA >= 0 || cB >= 0) {
          RawFace f{{pc(pp[0], d0[0]), pc(pp[1], d0[1]), pc(pp[2], d0[2]), pc(pp[3], d0[3])}, cA, cB, 3};
          faces_.push_back(std::move(f));
        }
        L = e;
      }
    }
}

// MRST's fan: area vector of a polygon around the mean of its corners.
static Vec3 fan_normal(const std::vector<Vec3>& x, Vec3* centre = nullptr) {
  Vec3 pc{0, 0, 0}, n{0, 0, 0};
  for (const Vec3& p : x)
    for (int d = 0; d < 3; ++d) pc[d] += p[d] / (double)x.size();
  for (size_t i = 0; i < x.size(); ++i) {
    const Vec3& a = x[i];
    const Vec3& b = x[(i + 1) % x.size()];
    const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double v[3] = {pc[0] - a[0], pc[1] - a[1], pc[2] - a[2]};
    n[0] += 0.5 * (u[1] * v[2] - u[2] * v[1]);
    n[1] += 0.5 * (u[2] * v[0] - u[0] * v[2]);
    n[2] += 0.5 * (u[0] * v[1] - u[1] * v[0]);
  }
  if (centre) *centre = pc;
  return n;
}

void Processor::run(CpGrid& g) {
  lateral(false);
  lateral(true);
  vertical();
  // cells: active with a face, numbered in Cartesian order
  const int ncart = nx_ * ny_ * nz_;
  std::vector<int> used(ncart, 0), id(ncart, -1);
  for (const RawFace& f : faces_) {
    if (f.c0 >= 0) used[f.c0] = 1;
    if (f.c1 >= 0) used[f.c1] = 1;
  }
  g.index_map.clear();
  for (int c = 0; c < ncart; ++c)
    if (used[c]) {
      id[c] = (int)g.index_map.size();
      g.index_map.push_back(c);
    }
  g.nc = (int)g.index_map.size();
  // cell centres (means of the corner points) and extents
  std::vector<Vec3> centre(g.nc);
  g.cell_dims.assign(3 * (size_t)g.nc, 0.0);
  for (int c = 0; c < g.nc; ++c) {
    const int cc = g.index_map[c];
    const int i = cc % nx_, j = (cc / nx_) % ny_, k = cc / (nx_ * ny_);
    Vec3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY}, m{0, 0, 0};
    for (int dk = 0; dk < 2; ++dk)
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const Vec3 x = point(pillar(i + di, j + dj), depth(2 * i + di, 2 * j + dj, 2 * k + dk));
          for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], x[d]);
            hi[d] = std::max(hi[d], x[d]);
            m[d] += x[d] / 8.0;
          }
        }
    centre[c] = m;
    for (int d = 0; d < 3; ++d) g.cell_dims[3 * (size_t)c + d] = hi[d] - lo[d];
  }
  // nodes: pillar points by (pillar, depth), crossing points by coordinates
  std::map<std::pair<int, double>, int> pkey;
  std::map<Vec3, int> xkey;
  std::vector<std::vector<const Corner*>> kept_corners;
  std::vector<RawFace*> kept;
  for (RawFace& f : faces_) {
    f.c0 = f.c0 >= 0 ? id[f.c0] : -1;
    f.c1 = f.c1 >= 0 ? id[f.c1] : -1;
    if (f.c0 < 0 && f.c1 < 0) continue;
    kept.push_back(&f);
  }
  for (RawFace* f : kept)
    for (const Corner& c : f->corners) {
      if (c.pillar >= 0) pkey.emplace(std::make_pair(c.pillar, c.d), 0);
      else xkey.emplace(c.x, 0);
    }
  int nn = 0;
  g.coords.clear();
  for (auto& kv : pkey) {
    kv.second = nn++;
    const Vec3 x = point(kv.first.first, kv.first.second);
    g.coords.insert(g.coords.end(), x.begin(), x.end());
  }
  for (auto& kv : xkey) {
    kv.second = nn++;
    g.coords.insert(g.coords.end(), kv.first.begin(), kv.first.end());
  }
  g.nn = nn;
  g.node_pos.assign(1, 0);
  g.face_nodes.clear();
  g.nb.clear();
  g.tag.clear();
  for (RawFace* f : kept) {
    std::vector<int> ids;
    std::vector<Vec3> xs;
    for (const Corner& c : f->corners) {
      const int n = c.pillar >= 0 ? pkey.at(std::make_pair(c.pillar, c.d)) : xkey.at(c.x);
      if (!ids.empty() && ids.back() == n) continue;
      ids.push_back(n);
      xs.push_back({g.coords[3 * (size_t)n], g.coords[3 * (size_t)n + 1], g.coords[3 * (size_t)n + 2]});
    }
    while (ids.size() > 1 && ids.front() == ids.back()) {
      ids.pop_back();
      xs.pop_back();
    }
    if (ids.size() < 3) continue;
    // orient from the first cell to the second
    Vec3 fc;
    const Vec3 n = fan_normal(xs, &fc);
    Vec3 dir;
    if (f->c0 >= 0)
      for (int d = 0; d < 3; ++d) dir[d] = fc[d] - centre[f->c0][d];
    else
      for (int d = 0; d < 3; ++d) dir[d] = centre[f->c1][d] - fc[d];
    if (n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2] < 0.0) std::reverse(ids.begin(), ids.end());
    g.face_nodes.insert(g.face_nodes.end(), ids.begin(), ids.end());
    g.node_pos.push_back((int)g.face_nodes.size());
    g.nb.push_back(f->c0);
    g.nb.push_back(f->c1);
    g.tag.push_back(f->tag);
  }
  g.nf = (int)g.tag.size();
  // faces of each cell
  g.face_pos.assign(g.nc + 1, 0);
  for (int f = 0; f < g.nf; ++f)
    for (int s = 0; s < 2; ++s)
      if (g.nb[2 * f + s] >= 0) ++g.face_pos[g.nb[2 * f + s] + 1];
  for (int c = 0; c < g.nc; ++c) g.face_pos[c + 1] += g.face_pos[c];
  g.cell_faces.assign(g.face_pos[g.nc], 0);
  std::vector<int> at(g.face_pos.begin(), g.face_pos.end() - 1);
  for (int f = 0; f < g.nf; ++f)
    for (int s = 0; s < 2; ++s)
      if (g.nb[2 * f + s] >= 0) g.cell_faces[at[g.nb[2 * f + s]]++] = f;
}

// ---------------------------------------------------------------------------
// geometry kernels
// ---------------------------------------------------------------------------
struct FanTri {
  double n[3], a, c[3];   // area vector, area, centroid of one fan triangle
};

__device__ __forceinline__ void face_center(int f, const int* __restrict__ node_pos, const int* __restrict__ fn,
                                            const double* __restrict__ x, double (&pc)[3]) {
  const int a = node_pos[f], b = node_pos[f + 1];
  pc[0] = pc[1] = pc[2] = 0.0;
  for (int q = a; q < b; ++q)
    for (int d = 0; d < 3; ++d) pc[d] += x[3 * (size_t)fn[q] + d];
  for (int d = 0; d < 3; ++d) pc[d] /= (double)(b - a);
}

__device__ __forceinline__ FanTri fan_tri(int q, int f, const int* __restrict__ node_pos, const int* __restrict__ fn,
                                          const double* __restrict__ x, const double (&pc)[3]) {
  const int a0 = node_pos[f], b0 = node_pos[f + 1];
  const int na = fn[q], nb = fn[q + 1 < b0 ? q + 1 : a0];
  const double* A = x + 3 * (size_t)na;
  const double* B = x + 3 * (size_t)nb;
  const double u[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
  const double v[3] = {pc[0] - A[0], pc[1] - A[1], pc[2] - A[2]};
  FanTri t;
  t.n[0] = 0.5 * (u[1] * v[2] - u[2] * v[1]);
  t.n[1] = 0.5 * (u[2] * v[0] - u[0] * v[2]);
  t.n[2] = 0.5 * (u[0] * v[1] - u[1] * v[0]);
  t.a = sqrt(t.n[0] * t.n[0] + t.n[1] * t.n[1] + t.n[2] * t.n[2]);
  for (int d = 0; d < 3; ++d) t.c[d] = (A[d] + B[d] + pc[d]) / 3.0;
  return t;
}

__global__ void k_cp_faces(int nf, const int* __restrict__ node_pos, const int* __restrict__ fn,
                           const double* __restrict__ x, double* __restrict__ fnormal, double* __restrict__ farea,
                           double* __restrict__ fcen) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double pc[3];
  face_center(f, node_pos, fn, x, pc);
  double n[3] = {0, 0, 0}, c[3] = {0, 0, 0}, area = 0.0;
  for (int q = node_pos[f]; q < node_pos[f + 1]; ++q) {
    const FanTri t = fan_tri(q, f, node_pos, fn, x, pc);
    for (int d = 0; d < 3; ++d) {
      n[d] += t.n[d];
      c[d] += t.a * t.c[d];
    }
    area += t.a;
  }
  for (int d = 0; d < 3; ++d) {
    fnormal[3 * (size_t)f + d] = n[d];
    fcen[3 * (size_t)f + d] = area > 0.0 ? c[d] / area : pc[d];
  }
  farea[f] = area;
}

__global__ void k_cp_cells(int nc, const int* __restrict__ face_pos, const int* __restrict__ cf,
                           const int* __restrict__ nb, const int* __restrict__ node_pos, const int* __restrict__ fn,
                           const double* __restrict__ x, const double* __restrict__ fnormal,
                           const double* __restrict__ fcen, double* __restrict__ vol, double* __restrict__ ccen) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int h0 = face_pos[c], h1 = face_pos[c + 1];
  double cc[3] = {0, 0, 0};
  for (int h = h0; h < h1; ++h)
    for (int d = 0; d < 3; ++d) cc[d] += fcen[3 * (size_t)cf[h] + d];
  for (int d = 0; d < 3; ++d) cc[d] /= (double)(h1 - h0);
  double v = 0.0, m[3] = {0, 0, 0};
  for (int h = h0; h < h1; ++h) {
    const int f = cf[h];
    const double sgn = nb[2 * f] == c ? 1.0 : -1.0;
    double pc[3];
    face_center(f, node_pos, fn, x, pc);
    const double* N = fnormal + 3 * (size_t)f;
    for (int q = node_pos[f]; q < node_pos[f + 1]; ++q) {
      const FanTri t = fan_tri(q, f, node_pos, fn, x, pc);
      const double dn = t.n[0] * N[0] + t.n[1] * N[1] + t.n[2] * N[2];
      const double s = (dn > 0.0 ? 1.0 : (dn < 0.0 ? -1.0 : 0.0)) * sgn;
      const double rel[3] = {t.c[0] - cc[0], t.c[1] - cc[1], t.c[2] - cc[2]};
      const double tv = (rel[0] * t.n[0] * s + rel[1] * t.n[1] * s + rel[2] * t.n[2] * s) / 3.0;
      v += tv;
      for (int d = 0; d < 3; ++d) m[d] += tv * 0.75 * rel[d];
    }
  }
  vol[c] = v;
  for (int d = 0; d < 3; ++d) ccen[3 * (size_t)c + d] = cc[d] + m[d] / v;
}

}  // namespace

void cp_process(int nx, int ny, int nz, const double* coord, const double* zcorn, const int* actnum, CpGrid& g) {
  PMX_CHECK(nx > 0 && ny > 0 && nz > 0, "corner-point dimensions must be positive");
  PMX_CHECK(coord && zcorn, "COORD and ZCORN are required");
  for (size_t q = 0; q < (size_t)8 * nx * ny * nz; ++q) PMX_CHECK(std::isfinite(zcorn[q]), "ZCORN is not finite");
  for (size_t q = 0; q < (size_t)6 * (nx + 1) * (ny + 1); ++q)
    PMX_CHECK(std::isfinite(coord[q]), "COORD is not finite");
  g = CpGrid();
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  Processor(nx, ny, nz, coord, zcorn, actnum).run(g);
}

void cp_geometry(const CpGrid& g, hipStream_t s, double* fnormal, double* farea, double* fcentroid, double* cvol,
                 double* ccentroid) {
  DArr<double> x;
  DArr<int> np, fn, nb, fp, cf;
  x.upload(g.coords, s);
  np.upload(g.node_pos, s);
  fn.upload(g.face_nodes, s);
  nb.upload(g.nb, s);
  fp.upload(g.face_pos, s);
  cf.upload(g.cell_faces, s);
  if (g.nf > 0) {
    hipLaunchKernelGGL(k_cp_faces, dim3(grid_for(g.nf)), dim3(256), 0, s, g.nf, np.p, fn.p, x.p, fnormal, farea,
                       fcentroid);
    PMX_HIP(hipGetLastError());
  }
  if (g.nc > 0) {
    hipLaunchKernelGGL(k_cp_cells, dim3(grid_for(g.nc)), dim3(256), 0, s, g.nc, fp.p, cf.p, nb.p, np.p, fn.p, x.p,
                       fnormal, fcentroid, cvol, ccentroid);
    PMX_HIP(hipGetLastError());
  }
  PMX_HIP(hipStreamSynchronize(s));
}

}  // namespace pmx
