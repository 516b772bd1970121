"""Node orders pinned by the reference's own vectors (CPU only).

The Gmsh and VTK node maps of the library are checked against the A.9 rows and DOLFINx's Gmsh P3
arrays that tetraloma-refcell/tests/node_maps.rs:11-52, 96-170 holds (copied here as data). The
oracle's A.5 node order is then pinned through those maps and the reference's Gmsh corpus
(tests/golden/gmsh, written by Gmsh 4.15.2): the corpus cells are affine images of their
reference cells, so the file's k-th node of a cell must sit at the affine image of the oracle's
node map[k] (direction A.9: internal[map[k]] = file[k]). The file is read by the plain-Python
tokenizer of tests/msh_sections.py, not by the library.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from gridcl_amd import gmsh, vtk
from gridcl_amd import meshes as M
from tests.msh_sections import read_ascii

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "gmsh")
IV, TR, QU, TE, HE, PR, PY = 1, 2, 3, 4, 5, 6, 7

# A.9 Gmsh rows (node_maps.rs:11-31).
A9_GMSH = [
    (IV, 1, [0, 1]),
    (IV, 2, [0, 1, 2]),
    (TR, 1, [0, 1, 2]),
    (TR, 2, [0, 1, 2, 3, 5, 4]),
    (QU, 1, [0, 1, 3, 2]),
    (QU, 2, [0, 1, 3, 2, 4, 6, 7, 5, 8]),
    (TE, 1, [0, 1, 2, 3]),
    (TE, 2, [0, 1, 2, 3, 4, 7, 5, 6, 9, 8]),
    (HE, 1, [0, 1, 3, 2, 4, 5, 7, 6]),
    (HE, 2, [0, 1, 3, 2, 4, 5, 7, 6, 8, 9, 10, 11, 12, 13, 15, 14, 16, 17, 18, 19,
             20, 21, 22, 23, 24, 25, 26]),
    (PR, 1, list(range(6))),
    (PR, 2, list(range(18))),
    (PY, 1, [0, 1, 3, 2, 4]),
    (PY, 2, [0, 1, 3, 2, 4, 5, 6, 7, 8, 9, 10, 12, 11, 13]),
]

# A.9 VTK Lagrange rows (node_maps.rs:33-52).
A9_VTK = [
    (IV, 1, [0, 1]),
    (IV, 2, [0, 1, 2]),
    (TR, 1, [0, 1, 2]),
    (TR, 2, [0, 1, 2, 3, 5, 4]),
    (QU, 1, [0, 1, 3, 2]),
    (QU, 2, [0, 1, 3, 2, 4, 6, 7, 5, 8]),
    (TE, 1, [0, 1, 2, 3]),
    (TE, 2, [0, 1, 2, 3, 4, 7, 5, 6, 8, 9]),
    (HE, 1, [0, 1, 3, 2, 4, 5, 7, 6]),
    (HE, 2, [0, 1, 3, 2, 4, 5, 7, 6, 8, 11, 13, 9, 16, 18, 19, 17, 10, 12, 15, 14,
             22, 23, 21, 24, 20, 25, 26]),
    (PR, 1, list(range(6))),
    (PR, 2, [0, 1, 2, 3, 4, 5, 6, 9, 7, 12, 14, 13, 8, 10, 11, 15, 17, 16]),
    (PY, 1, [0, 1, 3, 2, 4]),
]

# DOLFINx's Gmsh P3 arrays (node_maps.rs:96-130; DOLFINx cpp/dolfinx/io/cells.cpp).
DOLFINX_TET10 = [0, 1, 2, 3, 9, 6, 8, 7, 4, 5]
DOLFINX_TET20 = [0, 1, 2, 3, 14, 15, 8, 9, 13, 12, 11, 10, 5, 4, 7, 6, 19, 18, 17, 16]
DOLFINX_QUAD16 = [0, 1, 3, 2, 4, 5, 8, 9, 11, 10, 7, 6, 12, 13, 15, 14]
DOLFINX_HEX64 = [
    0, 1, 3, 2, 4, 5, 7, 6, 8, 9, 10, 11, 12, 13, 14, 15,
    16, 17, 19, 18, 22, 23, 20, 21, 24, 25, 26, 27, 28, 29, 31, 30,
    32, 34, 35, 33, 36, 37, 39, 38, 40, 42, 43, 41, 44, 45, 47, 46,
    49, 48, 50, 51, 52, 53, 55, 54, 56, 57, 59, 58, 60, 61, 63, 62,
]
DOLFINX_PRISM40 = [
    0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13,
    14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 37, 25, 26,
    28, 27, 29, 31, 32, 30, 33, 34, 36, 35, 38, 39,
]
DOLFINX_PYRAMID30 = [
    0, 1, 3, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14,
    16, 15, 19, 20, 17, 18, 25, 26, 27, 28, 21, 23, 24, 22, 29,
]


def basix_tet_to_internal(p: int, i: int) -> int:
    """A.11 for a tetrahedron of degree p (node_maps.rs:132-150): vertices and interior nodes
    keep their index, Basix edge e is edge 5 - e and Basix face f is face 3 - f, the order inside
    an entity's block unchanged."""
    per_edge = p - 1
    per_face = (p - 1) * (p - 2) // 2
    edges0 = 4
    faces0 = edges0 + 6 * per_edge
    interior0 = faces0 + 4 * per_face
    if i < edges0 or i >= interior0:
        return i
    if i < faces0:
        e, j = divmod(i - edges0, per_edge)
        return edges0 + (5 - e) * per_edge + j
    f, j = divmod(i - faces0, per_face)
    return faces0 + (3 - f) * per_face + j


GMSH_P3 = [
    (TE, 3, [basix_tet_to_internal(3, i) for i in DOLFINX_TET20]),
    (QU, 3, DOLFINX_QUAD16),
    (HE, 3, DOLFINX_HEX64),
    (PR, 3, DOLFINX_PRISM40),
    (PY, 3, DOLFINX_PYRAMID30),
]


def test_a11_reproduces_a9_tetrahedron_row():
    assert [basix_tet_to_internal(2, i) for i in DOLFINX_TET10] == A9_GMSH[7][2]


@pytest.mark.parametrize("cell,p,want", A9_GMSH + GMSH_P3)
def test_gmsh_map_matches_reference_vectors(cell, p, want):
    assert gmsh.gmsh_to_internal(cell, p).tolist() == want


@pytest.mark.parametrize("cell,p,want", A9_VTK)
def test_vtk_map_matches_reference_vectors(cell, p, want):
    assert vtk.vtk_to_internal(cell, p).tolist() == want


def test_interval_maps_are_identity():
    for p in range(1, 11):
        assert gmsh.gmsh_to_internal(IV, p).tolist() == list(range(p + 1))
        assert vtk.vtk_to_internal(IV, p).tolist() == list(range(p + 1))


# Gmsh element type -> (cell, degree) for the corpus files below (Gmsh's published numbering).
GMSH_TYPES = {1: (IV, 1), 8: (IV, 2), 26: (IV, 3), 2: (TR, 1), 9: (TR, 2), 21: (TR, 3),
              3: (QU, 1), 10: (QU, 2), 36: (QU, 3), 4: (TE, 1), 11: (TE, 2), 29: (TE, 3),
              5: (HE, 1), 12: (HE, 2), 92: (HE, 3), 6: (PR, 1), 13: (PR, 2), 90: (PR, 3),
              7: (PY, 1), 14: (PY, 2), 118: (PY, 3)}
# A.9: "Interval (any): identity" (node_maps.rs:85-94).
REFERENCE_MAPS = {(c, p): m for c, p, m in A9_GMSH + GMSH_P3 + [(IV, 3, [0, 1, 2, 3])]}
CORPUS = ["triangle_p1.msh", "triangle_p2.msh",
          "quadrilateral_p1.msh", "quadrilateral_p2.msh", "quadrilateral_p3.msh",
          "tetrahedron_p1.msh", "tetrahedron_p2.msh", "tetrahedron_p3.msh",
          "hexahedron_p1.msh", "hexahedron_p2.msh", "hexahedron_p3.msh",
          "prism_p1.msh", "prism_p2.msh", "prism_p3.msh",
          "hybrid_p1.msh", "hybrid_p2.msh", "hybrid_p3.msh", "interval_p3.msh"]


def _affine_fit(ref, phys):
    """Least-squares affine map ref -> phys and its worst residual."""
    a = np.hstack([ref, np.ones((len(ref), 1))])
    coef, *_ = np.linalg.lstsq(a, phys, rcond=None)
    return coef, np.abs(a @ coef - phys).max()


@pytest.mark.parametrize("name", CORPUS)
def test_oracle_nodes_at_corpus_positions(name):
    """For every corpus cell covered by a reference map: the affine image of the oracle's A.5
    nodes, read through the map, is the file's node list. Swapping two edges, two faces or the
    direction of an edge in the oracle's layout moves some node and fails here."""
    sec = read_ascii(os.path.join(GOLDEN, name))
    coords = {}
    for _, _, tags, xyz in sec.node_blocks:
        for t, x in zip(tags.tolist(), xyz):
            coords[t] = x
    checked = 0
    for _, _, ty, _, rows in sec.element_blocks:
        if ty not in GMSH_TYPES:
            continue
        cell, p = GMSH_TYPES[ty]
        if (cell, p) not in REFERENCE_MAPS:
            continue
        gmap = REFERENCE_MAPS[(cell, p)]
        ref = O.lagrange_points(cell, p)
        nv = {IV: 2, TR: 3, QU: 4, TE: 4, HE: 8, PR: 6, PY: 5}[cell]
        assert rows.shape[1] == len(ref) == len(gmap)
        for row in rows:
            filepts = np.array([coords[t] for t in row.tolist()])
            internal = np.empty_like(filepts)
            internal[gmap] = filepts
            coef, res_v = _affine_fit(ref[:nv], internal[:nv])
            scale = np.abs(internal[:nv] - internal[0]).max()
            if res_v > 1e-12 * scale:
                continue        # not an affine image (no corpus cell of the list is)
            image = np.hstack([ref, np.ones((len(ref), 1))]) @ coef
            assert np.abs(image - internal).max() <= 1e-9 * scale, (name, ty, row[:nv])
            checked += 1
    assert checked > 0, name


def test_library_nodes_equal_oracle_nodes():
    """The library's A.5 layout (refcell.cpp) equals the oracle's for every cell and degree the
    corpus test pins, and beyond it (to degree 5; pyramids 4)."""
    for cell, pmax in ((IV, 5), (TR, 5), (QU, 5), (TE, 5), (HE, 5), (PR, 5), (PY, 4)):
        for p in range(1, pmax + 1):
            assert np.array_equal(M.reference_nodes(cell, p), O.lagrange_points(cell, p)), (cell, p)
