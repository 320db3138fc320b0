"""Corner-point grids (ECLIPSE SPECGRID / COORD / ZCORN / ACTNUM) for the
device models: MRST's processGRDECL topology (host, C++) and computeGeometry
(device), then computeTrans (device) and the TPFA connection list the oil-water
and black-oil models take as ``N``, ``T``.

Reference: pymrst/gridprocessing/process_grdecl.py:352-580 (process_grdecl),
pymrst/gridprocessing/compute_geometry.py:211-325 (the 3D geometry),
pymrst/solvers/compute_trans.py:11-29 and
pymrst/solvers/get_face_transmissibility.py:12-27.

Faults (pillar pairs whose cell stacks do not match corner for corner) are
split into the overlaps of the active cells and gaps on either side, so a
faulted deck's connection list holds the cross-fault connections with
their overlap areas; zero-thickness layers join the cells above and below;
inactive and pinched cells are dropped.  Not processed: the PINCH keyword's
non-neighbour connections, MULTFLT/MULTX style multipliers, and
disconnected components (one grid is returned).
"""
import ctypes
from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

from . import _lib


@dataclass
class CornerPointGrid:
    """A processed corner-point grid in MRST's layout: faces.neighbors
    (-1 outside; the area vector points from the first cell to the second),
    faces.nodePos / faces.nodes, cells.facePos / cells.faces, cells.indexMap,
    and the geometry of computeGeometry."""
    cartDims: tuple
    num_cells: int
    num_faces: int
    num_nodes: int
    index_map: np.ndarray       # (nc,) Cartesian index i + nx (j + ny k)
    neighbors: np.ndarray       # (nf, 2)
    face_tag: np.ndarray        # (nf,) 1 i-, 2 j-, 3 k-face; +3 on a fault
    node_pos: np.ndarray        # (nf + 1,)
    face_nodes: np.ndarray
    node_coords: np.ndarray     # (nn, 3)
    face_pos: np.ndarray        # (nc + 1,)
    cell_faces: np.ndarray
    cell_dims: np.ndarray       # (nc, 3) extents of each cell's corner points
    cell_volumes: np.ndarray
    cell_centroids: np.ndarray  # (nc, 3)
    face_areas: np.ndarray
    face_normals: np.ndarray    # (nf, 3) area vectors
    face_centroids: np.ndarray  # (nf, 3)

    @property
    def coords(self):
        """Node coordinates per axis (as the tensor grids hold theirs)."""
        return tuple(self.node_coords[:, d] for d in range(3))

    @property
    def cells_num(self):
        return self.num_cells


def _corner_point_arrays(grdecl: dict):
    nx, ny, nz = (int(v) for v in grdecl["cartDims"])
    coord = np.ascontiguousarray(np.asarray(grdecl["COORD"], float).ravel())
    zcorn = np.ascontiguousarray(np.asarray(grdecl["ZCORN"], float).ravel())
    if coord.size != 6 * (nx + 1) * (ny + 1):
        raise ValueError(f"COORD holds {coord.size} values, {6 * (nx + 1) * (ny + 1)} expected")
    if zcorn.size != 8 * nx * ny * nz:
        raise ValueError(f"ZCORN holds {zcorn.size} values, {8 * nx * ny * nz} expected")
    act = grdecl.get("ACTNUM")
    if act is not None:
        act = np.ascontiguousarray((np.asarray(act).ravel() != 0).astype(np.int32))
        if act.size != nx * ny * nz:
            raise ValueError(f"ACTNUM holds {act.size} values, {nx * ny * nz} expected")
    return nx, ny, nz, coord, zcorn, act


def process_grdecl(grdecl: dict, device=None) -> CornerPointGrid:
    """processGRDECL + computeGeometry: the topology on the host (pmx_cp_process),
    the geometry on the GPU (pmx_cp_geometry).  ``grdecl`` holds cartDims,
    COORD, ZCORN and optionally ACTNUM, in the deck's order."""
    if not torch.cuda.is_available():
        raise _lib.PmxError("process_grdecl's geometry runs on the GPU (no CPU fallback)")
    lib = _lib.load()
    P = _lib.ptr
    nx, ny, nz, coord, zcorn, act = _corner_point_arrays(grdecl)
    h = ctypes.c_void_p()
    _lib.check(lib.pmx_cp_process(nx, ny, nz, P(coord), P(zcorn), P(act) if act is not None else None,
                                  ctypes.byref(h)), "pmx_cp_process")
    try:
        sizes = np.zeros(4, np.int32)
        _lib.check(lib.pmx_cp_sizes(h, P(sizes)), "pmx_cp_sizes")
        nc, nf, nn, ne = (int(v) for v in sizes)
        index_map = np.zeros(nc, np.int32)
        nb = np.zeros(2 * nf, np.int32)
        tag = np.zeros(nf, np.int32)
        node_pos = np.zeros(nf + 1, np.int32)
        face_nodes = np.zeros(ne, np.int32)
        xyz = np.zeros(3 * nn)
        face_pos = np.zeros(nc + 1, np.int32)
        # cells.faces holds one entry per face side with a cell
        cell_faces = np.zeros(2 * nf, np.int32)
        cell_dims = np.zeros(3 * nc)
        _lib.check(lib.pmx_cp_topology(h, P(index_map), P(nb), P(tag), P(node_pos), P(face_nodes), P(xyz),
                                       P(face_pos), P(cell_faces), P(cell_dims)), "pmx_cp_topology")
        dev = torch.device("cuda") if device is None else torch.device(device)
        f64 = dict(dtype=torch.float64, device=dev)
        fn, fa, fc = torch.empty(3 * nf, **f64), torch.empty(nf, **f64), torch.empty(3 * nf, **f64)
        cv, cc = torch.empty(nc, **f64), torch.empty(3 * nc, **f64)
        with torch.cuda.device(dev):
            stream = torch.cuda.current_stream(dev)
            _lib.check(lib.pmx_cp_geometry(h, dev.index or 0, ctypes.c_void_p(stream.cuda_stream), P(fn), P(fa),
                                           P(fc), P(cv), P(cc)), "pmx_cp_geometry")
    finally:
        lib.pmx_cp_destroy(h)
    nhf = int(face_pos[-1])
    return CornerPointGrid(
        cartDims=(nx, ny, nz), num_cells=nc, num_faces=nf, num_nodes=nn, index_map=index_map.astype(np.int64),
        neighbors=nb.reshape(nf, 2).astype(np.int64), face_tag=tag.astype(np.int64),
        node_pos=node_pos.astype(np.int64), face_nodes=face_nodes.astype(np.int64), node_coords=xyz.reshape(nn, 3),
        face_pos=face_pos.astype(np.int64), cell_faces=cell_faces[:nhf].astype(np.int64),
        cell_dims=cell_dims.reshape(nc, 3), cell_volumes=cv.cpu().numpy(),
        cell_centroids=cc.cpu().numpy().reshape(nc, 3), face_areas=fa.cpu().numpy(),
        face_normals=fn.cpu().numpy().reshape(nf, 3), face_centroids=fc.cpu().numpy().reshape(nf, 3))


def compute_trans(g: CornerPointGrid, rock, device=None) -> np.ndarray:
    """computeTrans on the GPU (pmx_compute_trans): the one-sided
    transmissibility of every half-face, in cells.faces order, for a
    diagonal (nc x 1 or 3) or full (nc x 6) permeability of the active
    cells."""
    lib = _lib.load()
    P = _lib.ptr
    dev = torch.device("cuda") if device is None else torch.device(device)
    perm = np.asarray(rock.perm, float)
    if perm.ndim == 1:
        perm = perm[:, None]
    if perm.shape[0] != g.num_cells or perm.shape[1] not in (1, 3, 6):
        raise ValueError("permeability must have one row (1, 3 or 6 values) per active cell")
    h = ctypes.c_void_p()
    with torch.cuda.device(dev):
        stream = torch.cuda.current_stream(dev)
        _lib.check(lib.pmx_incomp_create(dev.index or 0, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h)),
                   "pmx_incomp_create")
        try:
            fp = np.ascontiguousarray(g.face_pos, np.int32)
            cf = np.ascontiguousarray(g.cell_faces, np.int32)
            nb = np.ascontiguousarray(g.neighbors.ravel(), np.int32)
            cc = np.ascontiguousarray(g.cell_centroids.ravel(), float)
            fc = np.ascontiguousarray(g.face_centroids.ravel(), float)
            _lib.check(lib.pmx_incomp_set_grid(h, g.num_cells, g.num_faces, P(fp), P(cf), P(nb), P(cc), P(fc)),
                       "pmx_incomp_set_grid")
            fn = torch.as_tensor(np.ascontiguousarray(g.face_normals.ravel()), dtype=torch.float64, device=dev)
            k = torch.as_tensor(np.ascontiguousarray(perm.ravel()), dtype=torch.float64, device=dev)
            T = torch.empty(cf.size, dtype=torch.float64, device=dev)
            _lib.check(lib.pmx_compute_trans(h, P(fn), perm.shape[1], P(k), P(T)), "pmx_compute_trans")
            return T.cpu().numpy()
        finally:
            lib.pmx_incomp_destroy(h)


def tpfa_connections(g: CornerPointGrid, t_half: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """The internal connections N (n_if, 2) and their harmonic
    transmissibilities T = 1 / sum(1 / T_hf) (get_face_transmissibility.py:
    12-27, setup_operators_tpfa.py:98-110), in face order; fault overlaps
    are connections like any other face."""
    inv = np.zeros(g.num_faces)
    np.add.at(inv, g.cell_faces, 1.0 / np.asarray(t_half, float))
    internal = np.all(g.neighbors >= 0, axis=1)
    with np.errstate(divide="ignore"):
        T = 1.0 / inv[internal]
    return g.neighbors[internal].copy(), T
