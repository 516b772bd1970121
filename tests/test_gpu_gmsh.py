"""Gmsh files read into device geometries (the reader's cell blocks as the boundary's dofmaps):
the geometry path on the reference's Gmsh corpus (tests/golden/gmsh).

Pinned by: X at each cell's A.5 equispaced reference points equals the file's nodes read in A.5
order (FP64 1e-12; FP32 2e-6 relative to the coordinates' size), and the measure equals the sum of
the cells' volumes computed on the host from their vertices (every corpus cell is an affine image
of its reference cell, so its volume is the reference volume times sqrt(det(A^T A)) of the affine
map; FP64 1e-12 relative); a serendipity file has the measure of its complete-P2 twin."""
import json
import os

import numpy as np
import pytest

from gridcl_amd import api, gmsh
from gridcl_amd import meshes as M

pytestmark = pytest.mark.gpu

DIR = os.path.join(os.path.dirname(__file__), "golden", "gmsh")
FILES = [f for f in json.load(open(os.path.join(DIR, "manifest.json")))["files"]
         if "expect_error" not in f and not f["binary"]]
NV = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5}
REF_VOLUME = {1: 1.0, 2: 0.5, 3: 1.0, 4: 1.0 / 6.0, 5: 1.0, 6: 0.5, 7: 1.0 / 3.0}
AFFINE = ("interval", "triangle", "quadrilateral", "tetrahedron", "hexahedron", "prism", "hybrid",
          "two", "partitioned", "periodic", "square", "box")


def host_volume(m):
    """Sum over cells of the reference volume times the affine map's sqrt(det(A^T A))."""
    total = 0.0
    for blk in m.blocks:
        ref = M.reference_nodes(blk.cell, 1)
        a = np.concatenate([ref, np.ones((len(ref), 1))], axis=1)
        for row in blk.nodes:
            x = m.nodes[row[:NV[blk.cell]]]
            coef, *_ = np.linalg.lstsq(a, x, rcond=None)
            lin = coef[:-1].T  # [gdim][tdim]
            total += REF_VOLUME[blk.cell] * np.sqrt(np.linalg.det(lin.T @ lin))
    return total


@pytest.mark.parametrize("f", FILES, ids=[f["file"] for f in FILES])
def test_points_at_nodes_and_measure(cuda, f):
    import torch

    m = gmsh.read(os.path.join(DIR, f["file"]))
    g = m.geometry(torch.float64)
    for b, blk in enumerate(m.blocks):
        ref = M.reference_nodes(blk.cell, blk.degree)
        ps = api.PointSet.custom(blk.cell, ref, np.ones(len(ref)))
        x = api.physical_points_block(g, b, ps).to_host_cell_major()
        torch.cuda.synchronize()
        assert np.abs(x - m.nodes[blk.nodes]).max() <= 1e-12, b
    if f["file"].split("_")[0] in AFFINE:
        got = api.Mesh(g).measure(2)
        want = host_volume(m)
        assert abs(got - want) <= 1e-12 * abs(want)


def test_fp32_geometry_and_serendipity_twin(cuda):
    import torch

    m = gmsh.read(os.path.join(DIR, "hybrid_p2.msh"))
    g = m.geometry(torch.float32)
    scale = np.abs(m.nodes).max()
    for b, blk in enumerate(m.blocks):
        ref = M.reference_nodes(blk.cell, blk.degree)
        ps = api.PointSet.custom(blk.cell, ref, np.ones(len(ref)))
        x = api.physical_points_block(g, b, ps).to_host_cell_major()
        torch.cuda.synchronize()
        assert np.abs(x - m.nodes[blk.nodes]).max() <= 2e-6 * scale
    s = gmsh.read(os.path.join(DIR, "hybrid_s2.msh"))
    vs = api.Mesh(s.geometry(torch.float64)).measure(4)
    vt = api.Mesh(gmsh.read(os.path.join(DIR, "hybrid_p2.msh")).geometry(torch.float64)).measure(4)
    assert abs(vs - vt) <= 1e-13 * abs(vt)
