"""Cartesian grids, rock, transmissibilities and wells (host-side setup).

These build the inputs the TPFA path consumes, in MRST's ordering
(pymrst/gridprocessing/tensor_grid.py:118-206): cells x-fastest; faces are
the x-faces, then the y-faces, then the z-faces, each in x-fastest order;
each cell lists its faces W, E, S, N, T, B; ``neighbors[f]`` is (lower cell,
upper cell) with -1 outside.  Geometry is the exact tensor-grid geometry.
"""
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

# SI units (pymrst/utils/units/units.py:18-48)
MILLI, CENTI = 1e-3, 1e-2
DAY = 86400.0
BARSA = 1e5
POISE = 0.1
DARCY = (CENTI * 1.0) * (CENTI * POISE) / (101325.0 / CENTI)
NORM_GRAVITY = 9.80665


@dataclass
class Grid:
    cartDims: tuple
    coords: tuple                 # node coordinates per axis
    num_cells: int
    num_faces: int
    cell_faces: np.ndarray        # (nc, 6) face of each W,E,S,N,T,B side
    neighbors: np.ndarray         # (nf, 2)
    cell_centroids: np.ndarray    # (nc, 3)
    cell_volumes: np.ndarray      # (nc,)
    face_centroids: np.ndarray    # (nf, 3)
    face_normals: np.ndarray      # (nf, 3), from neighbors[:,0] to neighbors[:,1]
    cell_dims: np.ndarray         # (nc, 3) extents

    @property
    def cells_num(self):
        return self.num_cells


def cart_grid(dims: Sequence[int], physdims: Optional[Sequence[float]] = None) -> Grid:
    """cartGrid + computeGeometry for a box [0, L]^3 split into dims cells."""
    nx, ny, nz = (int(d) for d in dims)
    L = dims if physdims is None else physdims
    ax = [np.linspace(0.0, float(l), n + 1) for l, n in zip(L, (nx, ny, nz))]
    return tensor_grid(*ax)


def tensor_grid(x, y, z) -> Grid:
    x, y, z = (np.asarray(a, float) for a in (x, y, z))
    nx, ny, nz = x.size - 1, y.size - 1, z.size - 1
    nc = nx * ny * nz
    c = np.arange(nc)
    i, j, k = c % nx, (c // nx) % ny, c // (nx * ny)
    dx, dy, dz = np.diff(x), np.diff(y), np.diff(z)
    xm, ym, zm = x[:-1] + dx / 2, y[:-1] + dy / 2, z[:-1] + dz / 2
    nfx, nfy, nfz = (nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)
    fx = lambda ii, jj, kk: ii + (nx + 1) * (jj + ny * kk)
    fy = lambda ii, jj, kk: nfx + ii + nx * (jj + (ny + 1) * kk)
    fz = lambda ii, jj, kk: nfx + nfy + ii + nx * (jj + ny * kk)
    cell_faces = np.column_stack([fx(i, j, k), fx(i + 1, j, k), fy(i, j, k), fy(i, j + 1, k),
                                  fz(i, j, k), fz(i, j, k + 1)])
    nf = nfx + nfy + nfz
    nb = -np.ones((nf, 2), dtype=np.int64)
    nb[cell_faces[:, 1], 0] = c   # the cell below an E/N/B face is its first neighbour
    nb[cell_faces[:, 3], 0] = c
    nb[cell_faces[:, 5], 0] = c
    nb[cell_faces[:, 0], 1] = c
    nb[cell_faces[:, 2], 1] = c
    nb[cell_faces[:, 4], 1] = c
    fcent = np.zeros((nf, 3))
    fnorm = np.zeros((nf, 3))
    # x faces
    f = np.arange(nfx)
    fi, fj, fk = f % (nx + 1), (f // (nx + 1)) % ny, f // ((nx + 1) * ny)
    fcent[f] = np.column_stack([x[fi], ym[fj], zm[fk]])
    fnorm[f, 0] = dy[fj] * dz[fk]
    f = np.arange(nfy)
    fi, fj, fk = f % nx, (f // nx) % (ny + 1), f // (nx * (ny + 1))
    fcent[nfx + f] = np.column_stack([xm[fi], y[fj], zm[fk]])
    fnorm[nfx + f, 1] = dx[fi] * dz[fk]
    f = np.arange(nfz)
    fi, fj, fk = f % nx, (f // nx) % ny, f // (nx * ny)
    fcent[nfx + nfy + f] = np.column_stack([xm[fi], ym[fj], z[fk]])
    fnorm[nfx + nfy + f, 2] = dx[fi] * dy[fj]
    vol = dx[i] * dy[j] * dz[k]
    return Grid(cartDims=(nx, ny, nz), coords=(x, y, z), num_cells=nc, num_faces=nf,
                cell_faces=cell_faces, neighbors=nb,
                cell_centroids=np.column_stack([xm[i], ym[j], zm[k]]), cell_volumes=vol,
                face_centroids=fcent, face_normals=fnorm,
                cell_dims=np.column_stack([dx[i], dy[j], vol / (dx[i] * dy[j])]))


@dataclass
class Rock:
    perm: np.ndarray   # (nc, 1) isotropic or (nc, 3) diagonal
    poro: np.ndarray   # (nc,)


def make_rock(g: Grid, perm, poro) -> Rock:
    """makeRock (pymrst/params/rock/make_rock.py:14-60), diagonal tensors."""
    nc = g.num_cells
    k = np.asarray(perm, float)
    if k.ndim == 0:
        k = np.full((nc, 1), float(k))
    elif k.ndim == 1:
        k = k[:, None] if k.size == nc else np.tile(k[None, :], (nc, 1))
    elif k.shape[0] == 1:
        k = np.tile(k, (nc, 1))
    if k.shape[1] not in (1, 3):
        raise ValueError("permeability must be isotropic or diagonal (1 or 3 columns)")
    phi = np.asarray(poro, float).ravel()
    phi = np.full(nc, phi[0]) if phi.size == 1 else phi
    return Rock(perm=k, poro=phi)


def pore_volume(g: Grid, rock: Rock) -> np.ndarray:
    return rock.poro * g.cell_volumes


def compute_trans(g: Grid, rock: Rock) -> np.ndarray:
    """Half-face transmissibility, (nc*6,) in cell-major W,E,S,N,T,B order.

    computeTrans (pymrst/solvers/compute_trans.py:11-29) for a diagonal
    tensor: T = (c . K n)/|c|^2 with c the cell-to-face vector and n the
    outward area normal."""
    cf = g.cell_faces
    cc = g.cell_centroids
    kd = rock.perm if rock.perm.shape[1] == 3 else np.repeat(rock.perm, 3, axis=1)
    out = np.empty((g.num_cells, 6))
    for side in range(6):
        f = cf[:, side]
        cvec = g.face_centroids[f] - cc
        sgn = 1.0 if side % 2 == 1 else -1.0   # E/N/B faces point outward
        n = sgn * g.face_normals[f]
        out[:, side] = np.sum(cvec * kd * n, axis=1) / np.sum(cvec * cvec, axis=1)
    return np.abs(out).ravel()


def tpfa_connections(g: Grid, t_half: np.ndarray):
    """Internal connections N and harmonic transmissibilities T
    (get_face_transmissibility.py:12-27, setup_operators_tpfa.py:98-110)."""
    inv = np.zeros(g.num_faces)
    np.add.at(inv, g.cell_faces.ravel(), 1.0 / t_half)
    with np.errstate(divide="ignore"):
        T_all = 1.0 / inv
    internal = np.all(g.neighbors >= 0, axis=1)
    return g.neighbors[internal].copy(), T_all[internal].copy(), T_all, internal


def cartesian_lines(g: Grid, N: np.ndarray, T: np.ndarray):
    """Columns of a Cartesian grid along the axis with the largest total
    transmissibility, as (line_ptr, cells) for the line local stage of CPR.
    Lines are ordered by their first cell, so neighbouring lines hold
    neighbouring cells."""
    nx, ny, nz = g.cartDims
    strides = (1, nx, nx * ny)
    d = N[:, 1] - N[:, 0]
    totals = [float(T[d == s].sum()) for s in strides]
    axis = int(np.argmax(totals))
    n = (nx, ny, nz)[axis]
    idx = np.arange(g.num_cells).reshape((nz, ny, nx))   # [k, j, i]
    cols = np.moveaxis(idx, 2 - axis, -1).reshape(-1, n)  # one row per line, ascending along the axis
    order = np.argsort(cols[:, 0], kind="stable")
    cols = cols[order]
    ptr = np.arange(cols.shape[0] + 1, dtype=np.int32) * n
    return axis, ptr, np.ascontiguousarray(cols.ravel(), dtype=np.int32)


@dataclass
class Well:
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


def peaceman_wi(g: Grid, rock: Rock, cells, radius=0.1, skin=0.0) -> np.ndarray:
    """Vertical-well Peaceman index (compute_well_index.py:64-101, ip_tpf)."""
    d = g.cell_dims[cells]
    kd = rock.perm[cells] if rock.perm.shape[1] == 3 else np.repeat(rock.perm[cells], 3, axis=1)
    kx, ky = kd[:, 0], kd[:, 1]
    with np.errstate(divide="ignore", invalid="ignore"):
        kyx = np.where(kx > 0, ky / kx, 0.0)
        kxy = np.where(ky > 0, kx / ky, 0.0)
    re = 0.28 * np.sqrt(d[:, 0] ** 2 * np.sqrt(kyx) + d[:, 1] ** 2 * np.sqrt(kxy)) / (kyx ** 0.25 + kxy ** 0.25)
    return 2 * np.pi * d[:, 2] * np.sqrt(kx * ky) / (np.log(re / radius) + skin)


def add_well(wells: Optional[List[Well]], g: Grid, rock: Rock, cells, well_type="bhp", val=0.0,
             radius=0.1, compi=None, sign=0, name=None, gravity=None, lims=None, status=True) -> List[Well]:
    """addWell for vertical wells (pymrst/params/wells_and_bc/add_well.py:16-126)."""
    wells = list(wells or [])
    cells = np.asarray(cells, dtype=np.int64).ravel()
    if cells.size == 0 or cells.min() < 0 or cells.max() >= g.num_cells:
        raise IndexError("well completion cell index out of range")
    grav_on = gravity is not None and np.linalg.norm(np.ravel(gravity)) > 0
    ref = float(g.coords[2].min()) if grav_on else 0.0
    if well_type == "rate" and val != 0.0:
        sign = int(np.sign(val))
    wells.append(Well(cells=cells, type=well_type, val=float(val), WI=peaceman_wi(g, rock, cells, radius),
                      dZ=g.cell_centroids[cells, 2] - ref,
                      compi=np.array([1.0, 0.0, 0.0]) if compi is None else np.asarray(compi, float).ravel(),
                      sign=int(sign), name=name or f"W{len(wells) + 1}", lims=dict(lims) if lims else None,
                      status=bool(status)))
    return wells
