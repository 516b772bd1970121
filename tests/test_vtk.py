"""VTU output (tetraloma-io/src/vtk; tetraloma-refcell/src/node_maps/vtk.rs) on the host.

Pinned by: VTK's own files (tetraloma-io/tests/fixtures/vtk, one reference cell of each Lagrange
type and degree 1-6 written by VTK 9.7.1, copied under tests/golden/vtk): node k of each file sits
at the A.5 point that the library's VTK map sends it to, and a reference cell written by the
library lists the same points in the same order; the reference's doc-test examples (cell types,
the P2 quadrilateral's connectivity, linear output, the appended and base64 layouts); files read
back by an independent Python reader (XML, base64, raw appended data, zlib blocks) in every
encoding; the error variants of to_vtu."""
import json
import os

import numpy as np
import pytest

from gridcl_amd import api, vtk
from gridcl_amd import meshes as M

from tests import vtu_read

DIR = os.path.join(os.path.dirname(__file__), "golden", "vtk")
MANIFEST = json.load(open(os.path.join(DIR, "manifest.json")))
CELL = {68: M.INTERVAL, 69: M.TRIANGLE, 70: M.QUADRILATERAL, 71: M.TETRAHEDRON,
        72: M.HEXAHEDRON, 73: M.PRISM}
NV = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5}


@pytest.mark.parametrize("c", MANIFEST["cells"], ids=[c["file"] for c in MANIFEST["cells"]])
def test_node_order_matches_vtk_files(c, tmp_path):
    """VTK's file of one reference cell: its k-th connectivity point is the A.5 point
    vtk_to_internal(cell, p)[k]; the library writes the same cell with the same points in the
    same order and the same type."""
    f = vtu_read.read(os.path.join(DIR, c["file"]))
    cell, p = CELL[c["vtk_type"]], c["degree"]
    td = M.TDIM[cell]
    ref = M.reference_nodes(cell, p)
    vtk_pts = f["points"][f["connectivity"]][:, :td]
    assert len(vtk_pts) == c["num_points"] == len(ref)
    m = vtk.vtk_to_internal(cell, p)
    assert np.abs(ref[m] - vtk_pts).max() <= 1e-12
    mesh = M.HostMesh(td, ref, [M.HostBlock(cell, p, np.arange(len(ref), dtype=np.uint32)[None])])
    out = tmp_path / "cell.vtu"
    vtk.write_vtu(out, mesh)
    g = vtu_read.read(out)
    assert np.abs(g["points"][g["connectivity"]] - f["points"][f["connectivity"]]).max() <= 1e-15
    # Degree 1 is written as the linear type (cell_type.rs:41-60), VTK's file holds the Lagrange one.
    linear = {68: 3, 69: 5, 70: 9, 71: 10, 72: 12, 73: 13}[c["vtk_type"]]
    assert f["types"].tolist() == [c["vtk_type"]]
    assert g["types"].tolist() == [c["vtk_type"] if p >= 2 else linear]


def test_doc_examples(tmp_path):
    """cell_type.rs:31-38, convert.rs:73-98, write.rs:20-37."""
    assert vtk.vtk_cell_type(M.HEXAHEDRON, 1, True) == 12
    assert vtk.vtk_cell_type(M.HEXAHEDRON, 2, True) == 72
    assert vtk.vtk_cell_type(M.HEXAHEDRON, 2, False) == 12
    assert vtk.vtk_cell_type(M.PYRAMID, 3, True) == 14
    nodes = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0], [0, 0.5], [1, 0.5], [0.5, 1],
                      [0.5, 0.5]], dtype=float)
    mesh = M.HostMesh(2, nodes, [M.HostBlock(M.QUADRILATERAL, 2,
                                             np.arange(9, dtype=np.uint32)[None])])
    out = tmp_path / "q.vtu"
    assert vtk.write_vtu(out, mesh) == 0
    f = vtu_read.read(out)
    assert f["types"].tolist() == [70]
    assert f["connectivity"].tolist() == [0, 1, 3, 2, 4, 6, 7, 5, 8]
    assert f["offsets"].tolist() == [9]
    vtk.write_vtu(out, mesh, lagrange=False)
    f = vtu_read.read(out)
    assert f["types"].tolist() == [9] and f["connectivity"].tolist() == [0, 1, 3, 2]
    assert f["num_points"] == 9  # the non-corner nodes are written, unreferenced
    line = M.HostMesh(1, np.array([[0.0], [1.0]]),
                      [M.HostBlock(M.INTERVAL, 1, np.array([[0, 1]], dtype=np.uint32))])
    for base64, fmt in ((False, 'format="appended" offset="0"'), (True, 'format="binary"')):
        vtk.write_vtu(out, line, base64=base64)
        text = open(out, "rb").read().decode("latin-1")
        assert f'Name="Points" NumberOfComponents="3" {fmt}' in text
        assert ('<AppendedData encoding="raw">' in text) == (not base64)
        assert text.startswith('<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" '
                               'version="2.2" byte_order="LittleEndian" header_type="UInt64">')
    vtk.write_vtu(out, line, zlib=True)
    assert b'compressor="vtkZLibDataCompressor"' in open(out, "rb").read()


def mixed_mesh():
    """A hybrid mesh of every 3D cell type (P2 tetrahedra, Q2 hexahedra, P2 prisms, P2 pyramids)
    plus an unused node."""
    h = M.config4_mesh(seed=5, scale=0.0005)
    blocks = [b for b in h.blocks if M.TDIM[b.cell] == 3]
    out = []
    for b in blocks:
        if b.cell == M.PYRAMID:
            # Degree-2 pyramids from degree-1 ones: the writer uses their vertices only.
            k2 = 14
            cells = np.concatenate([b.cells, np.repeat(b.cells[:, :1], k2 - 5, axis=1)], axis=1)
            out.append(M.HostBlock(M.PYRAMID, 2, cells.astype(np.uint32)))
        else:
            out.append(b)
    nodes = np.concatenate([h.nodes, [[9.0, 9.0, 9.0]]])
    return M.HostMesh(3, nodes, out)


@pytest.mark.parametrize("base64", [False, True])
@pytest.mark.parametrize("zlib", [False, True])
@pytest.mark.parametrize("f32", [False, True])
def test_mixed_mesh_round_trip(tmp_path, base64, zlib, f32):
    """Every encoding reads back to the arrays to_vtu defines: points (Float32 rounded to nearest
    or Float64), each cell's nodes in VTK order, linear pyramids counted, cumulative offsets, the
    types, and point and cell data of several types."""
    mesh = mixed_mesh()
    n = mesh.num_nodes
    ncells = sum(b.num_cells for b in mesh.blocks)
    rng = np.random.default_rng(1)
    pd = {"speed": rng.standard_normal(n), "velocity": (rng.standard_normal(3 * n), 3),
          "id": np.arange(n, dtype=np.int64)}
    cd = {"quality": rng.standard_normal(ncells).astype(np.float32),
          "region": (np.arange(ncells) % 7).astype(np.uint8),
          "rank": np.zeros(ncells, dtype=np.int32)}
    out = tmp_path / "m.vtu"
    pyr = vtk.write_vtu(out, mesh, base64=base64, zlib=zlib, f32_points=f32, point_data=pd,
                        cell_data=cd)
    f = vtu_read.read(out)
    want_pts = np.asarray(mesh.nodes, dtype=np.float32 if f32 else np.float64)
    assert f["points_type"] == ("Float32" if f32 else "Float64")
    assert np.array_equal(f["points"], want_pts)
    conn, types, offs = [], [], []
    npyr = 0
    for b in mesh.blocks:
        linear = b.cell == M.PYRAMID and b.degree >= 2
        m = vtk.vtk_to_internal(b.cell, 1 if linear else b.degree)
        npyr += b.num_cells if linear else 0
        for row in b.cells:
            conn.extend(row[m].tolist())
            offs.append(len(conn))
            types.append(14 if linear else {4: 71, 5: 72, 6: 73}[b.cell])
    assert pyr == npyr > 0
    assert f["connectivity"].tolist() == conn
    assert f["offsets"].tolist() == offs and f["types"].tolist() == types
    assert np.array_equal(f["point_data"]["speed"][0], pd["speed"])
    assert f["point_data"]["velocity"][1] == 3
    assert np.array_equal(f["point_data"]["velocity"][0], pd["velocity"][0])
    assert np.array_equal(f["point_data"]["id"][0], pd["id"])
    assert np.array_equal(f["cell_data"]["quality"][0], cd["quality"])
    assert f["cell_data"]["region"][2] == "UInt8"
    assert np.array_equal(f["cell_data"]["region"][0], cd["region"])


def test_errors(tmp_path):
    out = tmp_path / "e.vtu"
    line = lambda cells, gdim=1, nodes=None: M.HostMesh(
        gdim, np.zeros((3, gdim)) if nodes is None else nodes,
        [M.HostBlock(M.INTERVAL, 1, np.asarray(cells, dtype=np.uint32))])
    with pytest.raises(api.DimensionMismatch):
        vtk.write_vtu(out, line([[0, 1]], gdim=4, nodes=np.zeros((3, 4))))
    with pytest.raises(api.IndexOutOfRange):
        vtk.write_vtu(out, line([[0, 3]]))
    with pytest.raises(api.LengthMismatch):
        vtk.write_vtu(out, line([[0, 1, 2]]))
    with pytest.raises(api.UnsupportedDegree):
        vtk.write_vtu(out, M.HostMesh(1, np.zeros((3, 1)),
                                      [M.HostBlock(M.INTERVAL, 0, np.zeros((1, 1), np.uint32))]))
    with pytest.raises(api.LengthMismatch):
        vtk.write_vtu(out, line([[0, 1]]), point_data={"x": np.zeros(2)})
    with pytest.raises(api.UnsupportedDegree):
        vtk.vtk_to_internal(M.PYRAMID, 2)
    with pytest.raises(api.Io):
        vtk.write_vtu(tmp_path / "no" / "dir" / "x.vtu", line([[0, 1]]))


def test_node_maps_are_permutations():
    for cell, pmax in ((M.INTERVAL, 10), (M.TRIANGLE, 10), (M.QUADRILATERAL, 10),
                       (M.TETRAHEDRON, 10), (M.HEXAHEDRON, 8), (M.PRISM, 8)):
        for p in range(1, pmax + 1):
            m = vtk.vtk_to_internal(cell, p)
            assert sorted(m.tolist()) == list(range(len(m)))
    assert vtk.vtk_to_internal(M.QUADRILATERAL, 1).tolist() == [0, 1, 3, 2]
    assert vtk.vtk_to_internal(M.HEXAHEDRON, 1).tolist() == [0, 1, 3, 2, 4, 5, 7, 6]
    assert vtk.vtk_to_internal(M.PYRAMID, 1).tolist() == [0, 1, 3, 2, 4]
