"""Corner-point grid checks (the oracle: test infrastructure only, never the
product path).

compute_geometry restates MRST's 3D computeGeometry
(pymrst/gridprocessing/compute_geometry.py:211-325) over a grid given by its
face nodes and cells' faces.  regular_grdecl is the corner-point deck of the
reference's own processGRDECL self-tests
(pymrst/gridprocessing/process_grdecl.py:585-603).  lateral_overlaps finds,
by brute force over every pair of cells on either side of every pillar pair,
the connections a corner-point grid with vertical pillars must have and
their exact overlap areas.
"""
import numpy as np


def compute_geometry(coords, node_pos, face_nodes, neighbors, face_pos, cell_faces):
    """(face normals, face areas, face centroids, cell volumes, cell
    centroids): faces as fans of triangles around the mean of their nodes,
    cells as tetrahedra from the mean of their face centroids
    (compute_geometry.py:211-308, _face_centroids :311-325)."""
    coords = np.asarray(coords, float)
    node_pos = np.asarray(node_pos, np.int64)
    face_nodes = np.asarray(face_nodes, np.int64)
    nf, nc = node_pos.size - 1, np.asarray(face_pos).size - 1
    num = np.diff(node_pos)
    face_no = np.repeat(np.arange(nf), num)
    pc = np.column_stack([np.bincount(face_no, coords[face_nodes, d], nf) for d in range(3)]) / num[:, None]
    nxt = np.arange(1, face_nodes.size + 1)
    nxt[node_pos[1:] - 1] = node_pos[:-1]
    a, b, p = coords[face_nodes], coords[face_nodes[nxt]], pc[face_no]
    sub_n = np.cross(b - a, p - a) / 2.0
    sub_a = np.sqrt((sub_n ** 2).sum(axis=1))
    sub_c = (a + b + p) / 3.0
    fnorm = np.column_stack([np.bincount(face_no, sub_n[:, d], nf) for d in range(3)])
    sub_s = np.sign((sub_n * fnorm[face_no]).sum(axis=1))
    farea = np.bincount(face_no, sub_a, nf)
    fcen = np.column_stack([np.bincount(face_no, sub_a * sub_c[:, d], nf) for d in range(3)])
    good = farea > 0
    fcen[good] /= farea[good, None]
    fcen[~good] = pc[~good]
    face_pos = np.asarray(face_pos, np.int64)
    cell_faces = np.asarray(cell_faces, np.int64)
    cell_no = np.repeat(np.arange(nc), np.diff(face_pos))
    cc = np.column_stack([np.bincount(cell_no, fcen[cell_faces, d], nc) for d in range(3)]) / \
        np.diff(face_pos)[:, None]
    vol = np.zeros(nc)
    mom = np.zeros((nc, 3))
    nb = np.asarray(neighbors, np.int64).reshape(-1, 2)
    for h, (c, f) in enumerate(zip(cell_no, cell_faces)):
        sgn = 1.0 if nb[f, 0] == c else -1.0
        for q in range(node_pos[f], node_pos[f + 1]):
            rel = sub_c[q] - cc[c]
            tv = rel @ (sub_n[q] * sub_s[q] * sgn) / 3.0
            vol[c] += tv
            mom[c] += tv * 0.75 * rel
    return fnorm, farea, fcen, vol, cc + mom / vol[:, None]


def regular_grdecl(nx, ny, nz, dx=100.0, dy=80.0, dz=5.0, fault_i=None, throw=0.0):
    """Vertical pillars on a dx x dy lattice, horizontal layers of
    thickness dz, the columns from fault_i on moved down by throw
    (process_grdecl.py:585-603)."""
    xp, yp = np.arange(nx + 1) * dx, np.arange(ny + 1) * dy
    coord = np.zeros(((nx + 1) * (ny + 1), 6))
    for j in range(ny + 1):
        for i in range(nx + 1):
            coord[i + (nx + 1) * j] = [xp[i], yp[j], 0.0, xp[i], yp[j], nz * dz]
    zc = np.zeros((2 * nx, 2 * ny, 2 * nz))
    for k in range(nz):
        zc[:, :, 2 * k] = k * dz
        zc[:, :, 2 * k + 1] = (k + 1) * dz
    if fault_i is not None:
        zc[2 * fault_i:, :, :] += throw
    return {"cartDims": (nx, ny, nz), "COORD": coord.ravel(), "ZCORN": zc.ravel(order="F")}


def _overlap_area(ta, ba, tb, bb, width):
    """Exact area of {max(ta, tb)(s) <= z <= min(ba, bb)(s), 0 <= s <= 1}
    times width, each argument a (value at s = 0, value at s = 1) pair of
    a line: the integrand is piecewise linear, exact by the trapezoid rule
    between its breakpoints."""
    lines = [ta, ba, tb, bb]
    brk = {0.0, 1.0}
    for i in range(4):
        for j in range(i + 1, 4):
            d0 = lines[i][0] - lines[j][0]
            d1 = lines[i][1] - lines[j][1]
            if d0 != d1:
                s = d0 / (d0 - d1)
                if 0.0 < s < 1.0:
                    brk.add(s)
    f = lambda L, s: L[0] + s * (L[1] - L[0])
    h = lambda s: max(0.0, min(f(ba, s), f(bb, s)) - max(f(ta, s), f(tb, s)))
    s = sorted(brk)
    return width * sum((s[q + 1] - s[q]) * (h(s[q]) + h(s[q + 1])) / 2 for q in range(len(s) - 1))


def lateral_overlaps(grdecl):
    """{(cell a, cell b): area} for every pair of active cells sharing a
    positive area of the vertical plane between two neighbouring columns
    (Cartesian cell indices, a < b), by brute force over all pairs.  The
    pillars must be vertical."""
    nx, ny, nz = grdecl["cartDims"]
    coord = np.asarray(grdecl["COORD"], float).reshape(-1, 6)
    if not np.allclose(coord[:, :2], coord[:, 3:5]):
        raise ValueError("lateral_overlaps needs vertical pillars")
    zc = np.asarray(grdecl["ZCORN"], float).reshape((2 * nx, 2 * ny, 2 * nz), order="F")
    act = np.ones(nx * ny * nz, bool) if grdecl.get("ACTNUM") is None else \
        np.asarray(grdecl["ACTNUM"]).ravel() != 0
    cart = lambda i, j, k: i + nx * (j + ny * k)
    pil = lambda i, j: coord[i + (nx + 1) * j, :2]
    out = {}
    for direction in ("i", "j"):
        rng = [(i, j) for i in range(nx - 1) for j in range(ny)] if direction == "i" else \
            [(i, j) for i in range(nx) for j in range(ny - 1)]
        for i, j in rng:
            if direction == "i":
                ia, ja, ib, jb, pa, pb = i, j, i + 1, j, (i + 1, j), (i + 1, j + 1)
                A = lambda k, t: (zc[2 * i + 1, 2 * j, 2 * k + t], zc[2 * i + 1, 2 * j + 1, 2 * k + t])
                B = lambda k, t: (zc[2 * i + 2, 2 * j, 2 * k + t], zc[2 * i + 2, 2 * j + 1, 2 * k + t])
            else:
                ia, ja, ib, jb, pa, pb = i, j, i, j + 1, (i, j + 1), (i + 1, j + 1)
                A = lambda k, t: (zc[2 * i, 2 * j + 1, 2 * k + t], zc[2 * i + 1, 2 * j + 1, 2 * k + t])
                B = lambda k, t: (zc[2 * i, 2 * j + 2, 2 * k + t], zc[2 * i + 1, 2 * j + 2, 2 * k + t])
            width = np.linalg.norm(pil(*pb) - pil(*pa))
            for ka in range(nz):
                ca = cart(ia, ja, ka)
                if not act[ca]:
                    continue
                for kb in range(nz):
                    cb = cart(ib, jb, kb)
                    if not act[cb]:
                        continue
                    area = _overlap_area(A(ka, 0), A(ka, 1), B(kb, 0), B(kb, 1), width)
                    if area > 1e-9 * width * 1.0:
                        out[(min(ca, cb), max(ca, cb))] = area
    return out
