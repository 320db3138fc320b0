"""The corner-point oracle (oracle/cpgrid.py) pinned on the CPU: its
computeGeometry restatement against closed forms of single hexahedra
(unit cube, a sheared cube, a wedge of linearly varying thickness) and the
divergence theorem, and its brute-force overlap areas against hand-computed
ones, each check with a planted mistake it must catch."""
import numpy as np
import pytest

from oracle import cpgrid as oc

from mutation import fails, planted


def hexahedra(corners):
    """Topology of separate hexahedral cells given their 8 corners each
    (order: x fastest, then y, then z, as a corner-point cell's), faces
    oriented outwards (no shared faces)."""
    coords, node_pos, face_nodes, nb, face_pos, cell_faces = [], [0], [], [], [0], []
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    for c, X in enumerate(corners):
        X = np.asarray(X, float)
        base = len(coords)
        coords.extend(X)
        centre = X.mean(axis=0)
        for q in quads:
            nodes = [base + v for v in q]
            P = X[list(q)]
            n = np.cross(P[1] - P[0], P[2] - P[0]) + np.cross(P[2] - P[0], P[3] - P[0])
            if n @ (P.mean(axis=0) - centre) < 0:      # outward from the cell: the cell is the first neighbour
                nodes = nodes[::-1]
            face_nodes.extend(nodes)
            node_pos.append(len(face_nodes))
            nb.append((c, -1))
            cell_faces.append(len(nb) - 1)
        face_pos.append(len(cell_faces))
    return np.array(coords), np.array(node_pos), np.array(face_nodes), np.array(nb), np.array(face_pos), \
        np.array(cell_faces)


def _box(f=lambda x, y, z: (x, y, z)):
    return [f(x, y, z) for z in (0.0, 1.0) for y in (0.0, 1.0) for x in (0.0, 1.0)]


def check_unit_cube():
    fn, fa, fc, vol, cc = oc.compute_geometry(*hexahedra([_box()]))
    assert vol[0] == pytest.approx(1.0, abs=1e-14)
    np.testing.assert_allclose(cc[0], [0.5, 0.5, 0.5], atol=1e-14)
    np.testing.assert_allclose(fa, 1.0, atol=1e-14)
    np.testing.assert_allclose(np.abs(fn).sum(axis=1), 1.0, atol=1e-14)


def check_sheared_cube():
    a = 0.7          # x shifted by a z: Cavalieri keeps the volume, the centroid moves by a/2
    fn, fa, fc, vol, cc = oc.compute_geometry(*hexahedra([_box(lambda x, y, z: (x + a * z, y, z))]))
    assert vol[0] == pytest.approx(1.0, abs=1e-13)
    np.testing.assert_allclose(cc[0], [0.5 + a / 2, 0.5, 0.5], atol=1e-13)


def check_wedge():
    c = 0.6          # thickness 1 + c x: volume 1 + c/2, centroid by moments
    fn, fa, fc, vol, cc = oc.compute_geometry(*hexahedra([_box(lambda x, y, z: (x, y, z * (1 + c * x)))]))
    V = 1 + c / 2
    xbar = (0.5 + c / 3) / V
    zbar = (0.5 + c / 2 + c * c / 6) / V      # integral of (1 + c x)^2 / 2 over the unit square
    assert vol[0] == pytest.approx(V, rel=1e-13)
    np.testing.assert_allclose(cc[0], [xbar, 0.5, zbar], rtol=1e-13)


def check_closed_cells():
    rng = np.random.default_rng(3)
    cells = [np.array(_box()) + rng.normal(0, 0.05, (8, 3)) for _ in range(5)]
    fn, fa, fc, vol, cc = oc.compute_geometry(*hexahedra(cells))
    net = fn.reshape(5, 6, 3).sum(axis=1)           # outward normals of each closed cell
    assert np.abs(net).max() < 1e-13
    # volume by the divergence theorem: V = sum over faces of (x_f . n_f) / 3
    v_div = (fc * fn).sum(axis=1).reshape(5, 6).sum(axis=1) / 3.0
    np.testing.assert_allclose(vol, v_div, rtol=1e-12)


def check_overlap_areas():
    # a face from depth 0 to 10 on both pillars against one from 4 to 14 on
    # the first pillar and -6 to 4 on the second: the overlap's height is
    # h(s) = min(10, 14 - 10 s) - max(0, 4 - 10 s) = 6 + 10 s on [0, 0.4]
    # and 14 - 10 s on [0.4, 1], width 2
    area = oc._overlap_area((0.0, 0.0), (10.0, 10.0), (4.0, -6.0), (14.0, 4.0), 2.0)
    exact = 2.0 * ((0.4 * 6 + 0.5 * 10 * 0.4 ** 2) + (14 * 0.6 - 5 * (1 - 0.16)))
    assert area == pytest.approx(exact, rel=1e-14)
    assert oc._overlap_area((0.0, 0.0), (1.0, 1.0), (1.0, 1.0), (2.0, 2.0), 1.0) == 0.0


CHECKS = {"unit_cube": check_unit_cube, "sheared": check_sheared_cube, "wedge": check_wedge,
          "closed": check_closed_cells, "overlap": check_overlap_areas}


@pytest.mark.parametrize("name", list(CHECKS))
def test_corner_point_oracle_known_answer(name):
    CHECKS[name]()


PLANTS = [
    ("compute_geometry", ("mom[c] += tv * 0.75 * rel", "mom[c] += tv * 0.5 * rel"), "wedge"),
    ("compute_geometry", ("sub_n = np.cross(b - a, p - a) / 2.0", "sub_n = np.cross(b - a, p - a)"), "unit_cube"),
    ("compute_geometry", ("sub_c = (a + b + p) / 3.0", "sub_c = (a + b) / 2.0"), "wedge"),
    ("_overlap_area", ("if 0.0 < s < 1.0:", "if False:"), "overlap"),
]


@pytest.mark.parametrize("func,edit,check", PLANTS)
def test_corner_point_oracle_sees_a_planted_mistake(func, edit, check):
    with planted(oc, func, edit):
        assert fails(CHECKS[check]), f"planted mistake not caught: {edit[0]!r}"
