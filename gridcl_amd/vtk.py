"""VTU output of meshes: VTK Lagrange cells written by the library's host writer.

Mirrors ``tetraloma_io::vtk::{to_vtu, VtuFile::write_to, vtk_cell_type}`` (tetraloma-io/src/vtk/
convert.rs:116, write.rs:378, cell_type.rs:41) for host meshes and ``tetraloma::VtkWriter::vtu``
(tetraloma/src/vtk.rs:736) for device geometries on one rank: version-2.2 UnstructuredGrid
files, raw appended or base64 data, optionally zlib-compressed; Lagrange cells (68-73) with their
nodes in VTK's order, pyramids of degree >= 2 from their vertices; a GLL geometry's nodes
converted to equispaced nodes on the device first.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib
from . import api

_TYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1, np.dtype(np.int32): 2,
          np.dtype(np.uint32): 3, np.dtype(np.int64): 4, np.dtype(np.uint8): 5}
LINEAR, BASE64, ZLIB, F32_POINTS = 1, 2, 4, 8


def _check(rc: int) -> None:
    if rc:
        raise api._ERRORS.get(rc, api.GridError)(_lib.last_error())


def vtk_to_internal(cell: int, degree: int) -> np.ndarray:
    """map[k] = the A.5 index of VTK's k-th node (tetraloma-refcell/src/node_maps.rs:107)."""
    lib = _lib.load()
    n = ctypes.c_size_t()
    _check(lib.tg_vtk_to_internal(cell, degree, None, 0, ctypes.byref(n)))
    out = (ctypes.c_int32 * n.value)()
    _check(lib.tg_vtk_to_internal(cell, degree, out, n.value, ctypes.byref(n)))
    return np.frombuffer(out, dtype=np.int32).copy()


def vtk_cell_type(cell: int, degree: int, lagrange: bool = True) -> int:
    return int(_lib.load().tg_vtk_cell_type(cell, degree, int(lagrange)))


def _options(lagrange, base64, zlib, f32_points=False) -> int:
    return ((0 if lagrange else LINEAR) | (BASE64 if base64 else 0) | (ZLIB if zlib else 0)
            | (F32_POINTS if f32_points else 0))


def _arrays(data, keep, device: bool):
    """tg_vtu_array descriptors of {name: (values, components)} or {name: values}."""
    items = list((data or {}).items())
    arr = (_lib.VtuArray * max(1, len(items)))()
    for i, (name, v) in enumerate(items):
        comps = 1
        if isinstance(v, tuple):
            v, comps = v
        if device:
            import torch

            dt = {torch.float32: 0, torch.float64: 1, torch.int32: 2, torch.int64: 4,
                  torch.uint8: 5}[v.dtype]
            t = v.contiguous()
            keep.append(t)
            ptr, count = t.data_ptr(), t.numel()
        else:
            a = np.ascontiguousarray(v)
            dt = _TYPES[a.dtype]
            keep.append(a)
            ptr, count = a.ctypes.data, a.size
        name_b = name.encode()
        keep.append(name_b)
        arr[i] = _lib.VtuArray(name_b, dt, comps, ptr, count)
    return arr, len(items)


def write_vtu(path, mesh, lagrange: bool = True, base64: bool = False, zlib: bool = False,
              f32_points: bool = False, point_data=None, cell_data=None) -> int:
    """Writes a host mesh (gdim, nodes [n][gdim], blocks with cell, degree, cells [c][k] in A.5
    order) as a VTU file; returns the number of pyramids written from their vertices. Data
    arrays: {name: values} or {name: (values, components)}, values as numpy arrays."""
    keep = []
    nodes = np.ascontiguousarray(np.asarray(mesh.nodes, dtype=np.float64))
    blocks = (_lib.VtuBlock * max(1, len(mesh.blocks)))()
    for i, b in enumerate(mesh.blocks):
        cells = np.ascontiguousarray(np.asarray(b.cells if hasattr(b, "cells") else b.nodes,
                                                dtype=np.uint32)).reshape(-1)
        keep.append(cells)
        blocks[i] = _lib.VtuBlock(int(b.cell), int(b.degree), cells.ctypes.data, cells.size)
    pd, npd = _arrays(point_data, keep, False)
    cd, ncd = _arrays(cell_data, keep, False)
    pyr = ctypes.c_uint64()
    _check(_lib.load().tg_vtu_write_mesh(
        os.fsencode(path), int(mesh.gdim), nodes.ctypes.data, nodes.shape[0], blocks,
        len(mesh.blocks), _options(lagrange, base64, zlib, f32_points), pd, npd, cd, ncd,
        ctypes.byref(pyr)))
    return int(pyr.value)


def write_geometry(path, g, lagrange: bool = True, base64: bool = False, zlib: bool = False,
                   point_data=None, cell_data=None, stream=None) -> int:
    """VtkWriter::vtu(path).write(mesh) for a device geometry (api.Geometry): equispaced nodes
    plus the origin as points (Float32 for FP32 geometries), cells block by block; data arrays
    are device tensors indexed by node and by flat cell index. Synchronises the stream."""
    import torch

    keep = []
    pd, npd = _arrays(point_data, keep, True)
    cd, ncd = _arrays(cell_data, keep, True)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    pyr = ctypes.c_uint64()
    _check(_lib.load().tg_vtu_write_geometry(g._h, os.fsencode(path),
                                             _options(lagrange, base64, zlib), pd, npd, cd, ncd,
                                             ctypes.byref(pyr), s))
    return int(pyr.value)
