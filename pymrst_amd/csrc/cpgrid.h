// Corner-point grids (ECLIPSE SPECGRID / COORD / ZCORN / ACTNUM): the
// topology of MRST's processGRDECL on the host, the geometry of
// computeGeometry on the device (see cpgrid.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "linsys.h"
#include "pmx_internal.h"

namespace pmx {

struct CpGrid {
  int nx = 0, ny = 0, nz = 0;
  int nc = 0, nf = 0, nn = 0;
  std::vector<double> coords;              // 3 per node
  std::vector<int> node_pos, face_nodes;   // nodes of each face (CSR), ordered so the face's area
                                           // vector points from nb[2f] to nb[2f+1]
  std::vector<int> nb;                     // 2 per face: the cells on either side, -1 outside
  std::vector<int> tag;                    // per face: 1 i-face, 2 j-face, 3 k-face; +3 on a fault
  std::vector<int> index_map;              // cell -> Cartesian index i + nx (j + ny k)
  std::vector<int> face_pos, cell_faces;   // faces of each cell (CSR, ascending)
  std::vector<double> cell_dims;           // 3 per cell: extents of its corner points
};

// processGRDECL's topology: coord[6 (nx+1)(ny+1)] pillar end points,
// zcorn[8 nx ny nz] corner depths, actnum[nx ny nz] (null: all active).
void cp_process(int nx, int ny, int nz, const double* coord, const double* zcorn, const int* actnum, CpGrid& g);

// computeGeometry on the device for a processed grid: per face the area
// vector (3), area and centroid (3); per cell the volume and centroid (3).
// All outputs are device arrays.
void cp_geometry(const CpGrid& g, hipStream_t s, double* fnormal, double* farea, double* fcentroid, double* cvol,
                 double* ccentroid);

}  // namespace pmx
