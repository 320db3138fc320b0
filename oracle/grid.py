"""Cartesian grid, geometry, computeTrans, face transmissibility, rock and
Peaceman wells -- oracle restatement (test infrastructure only).

Ordering follows MRST's tensor grid exactly
(pymrst/gridprocessing/tensor_grid.py:118-206): cells x-fastest; faces are
all x-faces, then y-faces, then z-faces, each block in Fortran order; each
cell lists its six faces with tags W, E, S, N, T, B; ``neighbors[f] =
(left/lower cell, right/upper cell)`` with -1 outside.  Geometry is the
analytic tensor-grid geometry, which is what compute_geometry
(pymrst/gridprocessing/compute_geometry.py:211-273) yields for planar,
axis-aligned faces (normals point from neighbors[:,0] to neighbors[:,1],
magnitude = face area).
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# pymrst/utils/units/units.py:18-48
MILLI, CENTI = 1e-3, 1e-2
DAY = 86400.0
BARSA = 1e5
POISE = 0.1
ATM = 101325.0
DARCY = (CENTI / 1.0) * (CENTI * POISE) / (ATM / CENTI)
NORM_GRAVITY = 9.80665  # pymrst/params/gravity_const.py:12


@dataclass
class CartGrid:
    dims: tuple
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    num_cells: int = 0
    num_faces: int = 0
    cell_faces: np.ndarray = None      # (6*nc,)
    cell_tags: np.ndarray = None       # (6*nc,)
    neighbors: np.ndarray = None       # (nf, 2)
    cell_centroids: np.ndarray = None  # (nc, 3)
    cell_volumes: np.ndarray = None    # (nc,)
    face_centroids: np.ndarray = None  # (nf, 3)
    face_normals: np.ndarray = None    # (nf, 3)
    face_areas: np.ndarray = None      # (nf,)
    cell_dx: np.ndarray = None         # (nc, 3) bounding-box extents

    def cell_no(self):
        return np.repeat(np.arange(self.num_cells), 6)


def cart_grid(celldim, physdims=None) -> CartGrid:
    """pymrst/gridprocessing/tensor_grid.py:44-63 (+ compute_geometry)."""
    celldim = tuple(int(d) for d in celldim)
    if physdims is None:
        physdims = celldim
    x, y, z = (np.linspace(0.0, float(p), d + 1) for p, d in zip(physdims, celldim))
    return tensor_grid(x, y, z)


def tensor_grid(x, y, z) -> CartGrid:
    """pymrst/gridprocessing/tensor_grid.py:118-206 for the 3-D case."""
    sx, sy, sz = x.size - 1, y.size - 1, z.size - 1
    nc = sx * sy * sz
    nfx, nfy, nfz = (sx + 1) * sy * sz, sx * (sy + 1) * sz, sx * sy * (sz + 1)
    nf = nfx + nfy + nfz
    F = lambda a: a.ravel(order="F")
    fx = np.arange(nfx).reshape((sx + 1, sy, sz), order="F")
    fy = nfx + np.arange(nfy).reshape((sx, sy + 1, sz), order="F")
    fz = nfx + nfy + np.arange(nfz).reshape((sx, sy, sz + 1), order="F")
    cell_faces = np.column_stack([F(fx[:sx]), F(fx[1:]), F(fy[:, :sy]), F(fy[:, 1:]),
                                  F(fz[:, :, :sz]), F(fz[:, :, 1:])]).ravel()
    tags = np.tile(np.arange(6), nc)
    c = -np.ones((sx + 2, sy + 2, sz + 2), dtype=np.int64)
    c[1:-1, 1:-1, 1:-1] = np.arange(nc).reshape((sx, sy, sz), order="F")
    nb = np.vstack([
        np.column_stack([F(c[:sx + 1, 1:sy + 1, 1:sz + 1]), F(c[1:, 1:sy + 1, 1:sz + 1])]),
        np.column_stack([F(c[1:sx + 1, :sy + 1, 1:sz + 1]), F(c[1:sx + 1, 1:, 1:sz + 1])]),
        np.column_stack([F(c[1:sx + 1, 1:sy + 1, :sz + 1]), F(c[1:sx + 1, 1:sy + 1, 1:])]),
    ])
    xm, ym, zm = 0.5 * (x[:-1] + x[1:]), 0.5 * (y[:-1] + y[1:]), 0.5 * (z[:-1] + z[1:])
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    I, J, K = np.meshgrid(np.arange(sx), np.arange(sy), np.arange(sz), indexing="ij")
    I, J, K = F(I), F(J), F(K)
    cc = np.column_stack([xm[I], ym[J], zm[K]])
    vol = dx[I] * dy[J] * dz[K]
    # face centroids / normals per block
    fc, fn, fa = [], [], []
    Ix, Jx, Kx = (F(a) for a in np.meshgrid(np.arange(sx + 1), np.arange(sy), np.arange(sz), indexing="ij"))
    fc.append(np.column_stack([x[Ix], ym[Jx], zm[Kx]]))
    a = dy[Jx] * dz[Kx]
    fa.append(a); fn.append(np.column_stack([a, 0 * a, 0 * a]))
    Iy, Jy, Ky = (F(a) for a in np.meshgrid(np.arange(sx), np.arange(sy + 1), np.arange(sz), indexing="ij"))
    fc.append(np.column_stack([xm[Iy], y[Jy], zm[Ky]]))
    a = dx[Iy] * dz[Ky]
    fa.append(a); fn.append(np.column_stack([0 * a, a, 0 * a]))
    Iz, Jz, Kz = (F(a) for a in np.meshgrid(np.arange(sx), np.arange(sy), np.arange(sz + 1), indexing="ij"))
    fc.append(np.column_stack([xm[Iz], ym[Jz], z[Kz]]))
    a = dx[Iz] * dy[Jz]
    fa.append(a); fn.append(np.column_stack([0 * a, 0 * a, a]))
    g = CartGrid(dims=(sx, sy, sz), x=x, y=y, z=z)
    g.num_cells, g.num_faces = nc, nf
    g.cell_faces, g.cell_tags, g.neighbors = cell_faces, tags, nb
    g.cell_centroids, g.cell_volumes = cc, vol
    g.face_centroids = np.vstack(fc)
    g.face_normals = np.vstack(fn)
    g.face_areas = np.concatenate(fa)
    g.cell_dx = np.column_stack([dx[I], dy[J], vol / (dx[I] * dy[J])])
    return g


@dataclass
class Rock:
    perm: np.ndarray   # (nc, 1) or (nc, 3)
    poro: np.ndarray   # (nc,)


def make_rock(g: CartGrid, perm, poro) -> Rock:
    """pymrst/params/rock/make_rock.py:14-60."""
    nc = g.num_cells
    perm = np.asarray(perm, dtype=float)
    if perm.ndim == 0:
        perm = np.full((nc, 1), float(perm))
    elif perm.ndim == 1:
        perm = perm.reshape(-1, 1) if perm.size == nc else np.repeat(perm.reshape(1, -1), nc, 0)
    elif perm.shape[0] == 1:
        perm = np.repeat(perm, nc, 0)
    poro = np.asarray(poro, dtype=float).ravel()
    if poro.size == 1:
        poro = np.full(nc, poro[0])
    return Rock(perm=perm, poro=poro)


def _perm_tensor(perm):
    """pymrst/params/rock/perm_tensor.py:37-60 (3-D): row-wise 9 entries."""
    nc, nk = perm.shape
    z = np.zeros(nc)
    if nk == 1:
        t6 = perm * np.array([1.0, 0.0, 0.0, 1.0, 0.0, 1.0])
    elif nk == 3:
        t6 = np.column_stack([perm[:, 0], z, z, perm[:, 1], z, perm[:, 2]])
    else:
        t6 = perm
    t9 = t6[:, [0, 1, 2, 1, 3, 4, 2, 4, 5]]
    row = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2])
    col = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2])
    return t9, row, col


def compute_trans(g: CartGrid, rock: Rock) -> np.ndarray:
    """One-sided transmissibility per half-face.

    pymrst/solvers/compute_trans.py:11-29: T = sum_ij c_i K_ij n_j / |c|^2,
    c = face centroid - cell centroid, n = outward area normal; the nine
    tensor terms are summed in the reference's order.
    """
    cell_no = g.cell_no()
    cf = g.cell_faces
    c = g.face_centroids[cf] - g.cell_centroids[cell_no]
    sgn = np.where(g.neighbors[cf, 0] == cell_no, 1.0, -1.0)
    n = sgn[:, None] * g.face_normals[cf]
    t9, row, col = _perm_tensor(rock.perm)
    T = np.zeros(cf.size)
    for k in range(9):
        T += c[:, row[k]] * t9[cell_no, k] * n[:, col[k]]
    T /= np.sum(c * c, axis=1)
    if np.any(T < 0):           # _fix_negative, compute_trans.py:32-42
        T = np.abs(T)
    return T


def face_trans(g: CartGrid, t_half: np.ndarray) -> np.ndarray:
    """Harmonic reduction per face, pymrst/solvers/get_face_transmissibility.py:12-27."""
    with np.errstate(divide="ignore"):
        return 1.0 / np.bincount(g.cell_faces, weights=1.0 / t_half, minlength=g.num_faces)


def internal_connections(g: CartGrid, t_half: np.ndarray):
    """N, T (internal faces) as setup_operators_tpfa builds them
    (pymrst/modules/ad_core/utils/setup_operators_tpfa.py:98-110)."""
    T_all = face_trans(g, t_half)
    i = np.all(g.neighbors >= 0, axis=1)
    return g.neighbors[i].copy(), T_all[i].copy(), T_all, i


@dataclass
class Well:
    """Fields used on the path (pymrst/params/wells_and_bc/types.py:42-72)."""
    cells: np.ndarray
    type: str
    val: float
    WI: np.ndarray
    dZ: np.ndarray
    compi: np.ndarray
    sign: int = 0
    name: str = ""
    lims: Optional[dict] = None      # limits a controlled well may switch to (types.py:42-72)
    status: bool = True              # an open well


def well_index(g: CartGrid, rock: Rock, cells, radius=0.1, skin=0.0, well_dir="z"):
    """Peaceman WI, pymrst/utils/compute_well_index.py:64-101 with the
    ip_tpf constant 0.14 and kh = ell*sqrt(k1 k2) (:151-162)."""
    cells = np.asarray(cells, dtype=np.int64)
    dims = g.cell_dx[cells]
    perm = rock.perm
    cols = {1: [0, 0, 0], 3: [0, 1, 2], 6: [0, 3, 5]}[perm.shape[1]]
    kd = perm[np.ix_(cells, cols)]
    ax = {"x": (1, 2, 0), "y": (0, 2, 1), "z": (0, 1, 2)}[well_dir]
    d1, d2, ell = dims[:, ax[0]], dims[:, ax[1]], dims[:, ax[2]]
    k1, k2 = kd[:, ax[0]], kd[:, ax[1]]
    k21 = np.where(k1 > 0, k2 / k1, 0.0)
    k12 = np.where(k2 > 0, k1 / k2, 0.0)
    re = 2 * 0.14 * np.sqrt(d1 ** 2 * np.sqrt(k21) + d2 ** 2 * np.sqrt(k12)) / (k21 ** 0.25 + k12 ** 0.25)
    kh = ell * np.sqrt(k1 * k2)
    radius = np.broadcast_to(np.asarray(radius, float), cells.shape)
    return 2 * np.pi * kh / (np.log(re / radius) + skin)


def add_well(wells: Optional[List[Well]], g: CartGrid, rock: Rock, cells, well_type="bhp",
             val=0.0, radius=0.1, compi=None, sign=0, gravity=None, name=None,
             ref_depth=None, well_dir="z", lims=None, status=True) -> List[Well]:
    """pymrst/params/wells_and_bc/add_well.py:16-126 (well_dir x, y or z).

    dZ = depth - ref_depth where depth is the cell centroid along the
    gravity direction (z when gravity is off) and ref_depth is the shallowest
    node (0 when gravity is off), add_well.py:94-96, 275-307.
    """
    wells = list(wells or [])
    cells = np.asarray(cells, dtype=np.int64).ravel()
    wi = well_index(g, rock, cells, radius, well_dir=well_dir)
    compi = np.array([1.0, 0.0, 0.0]) if compi is None else np.asarray(compi, float).ravel()
    depth = g.cell_centroids[cells, 2]
    if ref_depth is None:
        grav_on = gravity is not None and np.linalg.norm(np.ravel(gravity)) > 0
        ref_depth = float(g.z.min()) if grav_on else 0.0
    if well_type == "rate" and val != 0.0:
        sign = int(np.sign(val))
    wells.append(Well(cells=cells, type=well_type, val=float(val), WI=wi,
                      dZ=depth - ref_depth, compi=compi, sign=int(sign),
                      name=name or f"W{len(wells) + 1}", lims=dict(lims) if lims else None,
                      status=bool(status)))
    return wells
