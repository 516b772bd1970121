"""The C ABI without a GPU: the library loads, exports every declared symbol, and its host-side
checks return the reference's error variants before any device work (CPU only)."""
import ctypes
import os
import re

import numpy as np
import pytest

from gridcl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tetraloma_geom.h")

FAKE = 0x1000  # a non-null device address; none of these tests reaches a kernel launch


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding and header disagree"


def test_version_and_counts():
    lib = _lib.load()
    assert b"gfx950" in lib.tg_version()
    assert lib.tg_num_lagrange_nodes(4, 2) == 10
    assert lib.tg_num_lagrange_nodes(5, 3) == 64
    assert lib.tg_field_elems(20, 4, 3, 16) == 32 * 4 * 3  # field.rs:214-219 doc example
    assert lib.tg_field_elems(5, 3, 1, 4) == 8 * 3
    assert lib.tg_field_elems(5, 3, 1, 3) == 0


def _geometry(block, dtype=0, gdim=3, nodes=FAKE, num_nodes=10):
    lib = _lib.load()
    d = (_lib.BlockDesc * 1)()
    d[0].cell, d[0].degree, d[0].family = block[0], block[1], block[2]
    d[0].dofmap = FAKE
    d[0].dofmap_len, d[0].dofmap_stride = block[3], block[4]
    d[0].flags = None
    h = ctypes.c_void_p()
    org = (ctypes.c_double * 3)(0, 0, 0)
    rc = lib.tg_geometry_create(dtype, gdim, nodes, num_nodes, org, d, 1, ctypes.byref(h))
    return rc, h


def test_geometry_checks():
    """geometry.rs:106-150: stride must equal k and the length be a multiple of it."""
    from gridcl_amd import api

    rc, h = _geometry((4, 2, 0, 20, 10))
    assert rc == 0
    assert _lib.load().tg_geometry_block_cells(h, 0) == 2
    _lib.load().tg_geometry_destroy(h)
    rc, _ = _geometry((4, 2, 0, 20, 4))
    assert rc == api.LengthMismatch.code and "LengthMismatch" in _lib.last_error()
    rc, _ = _geometry((4, 2, 0, 21, 10))
    assert rc == api.LengthMismatch.code
    rc, _ = _geometry((4, 2, 0, 10, 10), num_nodes=0, nodes=None)
    assert rc == api.LengthMismatch.code
    rc, _ = _geometry((4, 0, 0, 4, 1))
    assert rc == api.UnsupportedDegree.code
    rc, _ = _geometry((4, 2, 1, 10, 10))
    assert rc == api.UnsupportedNodeFamily.code
    rc, _ = _geometry((4, 2, 0, 10, 10), dtype=5)
    assert rc == api.UnsupportedPrecision.code
    rc, _ = _geometry((4, 1, 0, 4, 4), gdim=2)
    assert rc == api.InvalidArgument.code  # a tetrahedron cannot live in 2D


def test_eval_argument_checks():
    from gridcl_amd import api

    lib = _lib.load()
    rc, g = _geometry((4, 2, 0, 20, 10))
    assert rc == 0
    hex_ps = api.PointSet.gauss(5, 2)
    tet_ps = api.PointSet.gauss(4, 2)
    out = _lib.Outputs()
    out.detj = FAKE
    # CellNotInRule: a point set of another cell type (rule.rs:163-172).
    assert lib.tg_eval_block(g, 0, hex_ps._h, 4, 16, 0, ctypes.byref(out), None) == 6
    assert "CellNotInRule" in _lib.last_error()
    # Bad tile, bad mask, missing output, block out of range.
    assert lib.tg_eval_block(g, 0, tet_ps._h, 4, 3, 0, ctypes.byref(out), None) == 8
    assert lib.tg_eval_block(g, 0, tet_ps._h, 0, 16, 0, ctypes.byref(out), None) == 8
    assert lib.tg_eval_block(g, 0, tet_ps._h, 1, 16, 0, ctypes.byref(out), None) == 8
    assert lib.tg_eval_block(g, 1, tet_ps._h, 4, 16, 0, ctypes.byref(out), None) == 8
    # Normals need codimension 1; facet normals a facet point set.
    out.normals = FAKE
    out.facet_normals = FAKE
    assert lib.tg_eval_block(g, 0, tet_ps._h, 32, 16, 0, ctypes.byref(out), None) == 8
    assert lib.tg_eval_block(g, 0, tet_ps._h, 16, 16, 0, ctypes.byref(out), None) == 8
    lib.tg_geometry_destroy(g)


def test_empty_launches_nothing():
    """An empty block returns OK before touching the device (layout.rs:509-511)."""
    from gridcl_amd import api

    lib = _lib.load()
    rc, g = _geometry((4, 1, 0, 0, 4))
    assert rc == 0
    out = _lib.Outputs()
    out.detj = FAKE
    ps = api.PointSet.tet_keast4()
    assert lib.tg_eval_block(g, 0, ps._h, 4, 16, 0, ctypes.byref(out), None) == 0
    empty = api.PointSet.custom(4, np.zeros((0, 3)), np.zeros(0))
    rc, g2 = _geometry((4, 1, 0, 8, 4))
    assert lib.tg_eval_block(g2, 0, empty._h, 4, 16, 0, ctypes.byref(out), None) == 0


def test_tabulation_errors():
    from gridcl_amd import api

    ps = api.PointSet.gauss(4, 2)
    with pytest.raises(api.UnsupportedNodeFamily):
        ps.tabulate(api.CoordinateElement(api.CellType.TETRAHEDRON, 2, api.NodeFamily.GLL))
    with pytest.raises(api.UnsupportedDegree):
        ps.tabulate(api.CoordinateElement(api.CellType.TETRAHEDRON, 0))
    with pytest.raises(api.CellNotInRule):
        ps.tabulate(api.CoordinateElement(api.CellType.HEXAHEDRON, 1))


def test_no_cpu_fallback_in_product():
    """The package never imports the oracle or computes factors on the host."""
    pkg = os.path.join(ROOT, "gridcl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text, f
