"""ctypes binding of the C ABI declared in include/tetraloma_geom.h.

The library is built in-tree (``gridcl_amd/lib/libtetraloma_geom.so``) by
``gridcl_amd.build.build_library``. There is no fallback: importing a module that needs the
library raises if it is missing, so no caller can silently run anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_int32, c_int64, c_size_t, c_uint32, c_uint64,
                    c_void_p)

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libtetraloma_geom.so")

# Every symbol the header declares, with (restype, argtypes).
SIGNATURES = {
    "tg_last_error_message": (c_char_p, []),
    "tg_version": (c_char_p, []),
    "tg_field_elems": (c_uint64, [c_uint64, c_uint64, c_uint32, c_uint32]),
    "tg_num_lagrange_nodes": (c_uint64, [c_int32, c_uint32]),
    "tg_lagrange_points": (
        c_int32,
        [c_int32, c_uint32, c_int32, POINTER(c_double), c_size_t, POINTER(c_size_t)],
    ),
    "tg_geometry_create": (
        c_int32,
        [c_int32, c_int32, c_void_p, c_uint64, POINTER(c_double), c_void_p, c_int32,
         POINTER(c_void_p)],
    ),
    "tg_geometry_destroy": (c_int32, [c_void_p]),
    "tg_geometry_num_blocks": (c_int32, [c_void_p]),
    "tg_geometry_block_cells": (c_uint64, [c_void_p, c_int32]),
    "tg_geometry_set_fixup_threshold": (c_int32, [c_void_p, c_double]),
    "tg_geometry_set_nodes_f64": (c_int32, [c_void_p, c_void_p]),
    "tg_pointset_create": (
        c_int32,
        [c_int32, POINTER(c_double), POINTER(c_double), c_uint64, POINTER(c_double),
         POINTER(c_void_p)],
    ),
    "tg_pointset_gauss": (c_int32, [c_int32, c_uint32, POINTER(c_void_p)]),
    "tg_pointset_facet_gauss": (c_int32, [c_int32, c_uint32, POINTER(c_void_p)]),
    "tg_pointset_tet_rule": (c_int32, [c_int32, POINTER(c_void_p)]),
    "tg_pointset_destroy": (c_int32, [c_void_p]),
    "tg_pointset_len": (c_uint64, [c_void_p]),
    "tg_pointset_get": (
        c_int32, [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                  POINTER(c_uint64)]
    ),
    "tg_pointset_tabulate": (
        c_int32, [c_void_p, c_uint32, c_int32, POINTER(c_double), POINTER(c_double)]
    ),
    "tg_eval_block": (
        c_int32, [c_void_p, c_int32, c_void_p, c_uint32, c_uint32, c_uint32, c_void_p, c_void_p]
    ),
    "tg_evector_build": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
    "tg_evector_elems": (c_uint64, [c_void_p, c_int32]),
    "tg_plan_create": (
        c_int32, [c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_void_p, POINTER(c_void_p)]
    ),
    "tg_plan_destroy": (c_int32, [c_void_p]),
    "tg_plan_info": (
        c_int32, [c_void_p, POINTER(c_uint32), POINTER(c_uint32), POINTER(c_int32)]
    ),
    "tg_plan_launch": (c_int32, [c_void_p, c_void_p]),
    "tg_plan_graph_launch": (c_int32, [c_void_p, c_void_p]),
    "tg_measure_block": (c_int32, [c_void_p, c_int32, c_void_p, c_uint32, c_void_p, c_void_p]),
    "tg_measure_block_noalloc": (c_int32, [c_void_p, c_int32, c_void_p, c_uint32, c_void_p,
                                           c_void_p]),
    "tg_equispaced_nodes": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "tg_pull_back_block": (
        c_int32, [c_void_p, c_int32, c_void_p, c_uint64, c_uint32, c_uint32, c_double, c_uint32,
                  c_void_p, c_void_p, c_void_p]
    ),
    "tg_functional_create": (c_int32, [c_void_p, c_int32, c_void_p, POINTER(c_void_p)]),
    "tg_functional_destroy": (c_int32, [c_void_p]),
    "tg_functional_forward": (c_int32, [c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "tg_functional_vjp": (
        c_int32, [c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p]
    ),
    "tg_gmsh_read_file": (c_int32, [c_char_p, c_int32, POINTER(c_void_p)]),
    "tg_gmsh_read_bytes": (c_int32, [c_char_p, c_uint64, c_int32, POINTER(c_void_p)]),
    "tg_gmsh_destroy": (c_int32, [c_void_p]),
    "tg_gmsh_last_error_detail": (c_int64, []),
    "tg_gmsh_get_info": (c_int32, [c_void_p, c_void_p]),
    "tg_gmsh_nodes": (c_void_p, [c_void_p]),
    "tg_gmsh_node_tags": (c_void_p, [c_void_p]),
    "tg_gmsh_get_block": (c_int32, [c_void_p, c_int32, c_void_p]),
    "tg_gmsh_get_tag_group": (c_int32, [c_void_p, c_int32, c_void_p]),
    "tg_gmsh_get_name": (
        c_int32, [c_void_p, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_char_p)]
    ),
    "tg_gmsh_get_element_block": (c_int32, [c_void_p, c_int32, c_void_p]),
    "tg_vtk_to_internal": (
        c_int32, [c_int32, c_uint32, POINTER(c_int32), c_size_t, POINTER(c_size_t)]
    ),
    "tg_vtk_cell_type": (c_int32, [c_int32, c_uint32, c_int32]),
    "tg_vtu_write_mesh": (
        c_int32, [c_char_p, c_int32, c_void_p, c_uint64, c_void_p, c_int32, c_uint32, c_void_p,
                  c_int32, c_void_p, c_int32, POINTER(c_uint64)]
    ),
    "tg_vtu_write_geometry": (
        c_int32, [c_void_p, c_char_p, c_uint32, c_void_p, c_int32, c_void_p, c_int32,
                  POINTER(c_uint64), c_void_p]
    ),
    "tg_gmsh_to_internal": (
        c_int32, [c_int32, c_uint32, POINTER(c_int32), c_size_t, POINTER(c_size_t)]
    ),
}


class BlockDesc(ctypes.Structure):
    """tg_block_desc."""

    _fields_ = [
        ("cell", c_int32),
        ("degree", c_uint32),
        ("family", c_int32),
        ("dofmap", c_void_p),
        ("dofmap_len", c_uint64),
        ("dofmap_stride", c_uint32),
        ("flags", c_void_p),
    ]


class Outputs(ctypes.Structure):
    """tg_outputs."""

    _fields_ = [
        ("points", c_void_p),
        ("jacobians", c_void_p),
        ("detj", c_void_p),
        ("pinv", c_void_p),
        ("facet_normals", c_void_p),
        ("normals", c_void_p),
    ]


class GmshInfo(ctypes.Structure):
    """tg_gmsh_info."""

    _fields_ = [
        ("tdim", c_int32),
        ("gdim", c_int32),
        ("num_nodes", c_uint64),
        ("num_blocks", c_int32),
        ("num_tag_groups", c_int32),
        ("num_names", c_int32),
        ("binary", c_int32),
        ("file_nodes", c_uint64),
        ("num_element_blocks", c_int32),
        ("partitioned", c_int32),
        ("num_partitions", c_uint64),
        ("num_periodic", c_uint64),
    ]


class GmshBlock(ctypes.Structure):
    """tg_gmsh_block."""

    _fields_ = [
        ("cell", c_int32),
        ("degree", c_uint32),
        ("num_cells", c_uint64),
        ("nodes_per_cell", c_uint32),
        ("nodes", c_void_p),
        ("tags", c_void_p),
        ("elementary", c_void_p),
        ("num_physical", c_uint64),
        ("physical", c_void_p),
    ]


class GmshTagGroup(ctypes.Structure):
    """tg_gmsh_tag_group."""

    _fields_ = [
        ("dim", c_int32),
        ("cell", c_int32),
        ("num_elements", c_uint64),
        ("corners_per_element", c_uint32),
        ("corners", c_void_p),
        ("tags", c_void_p),
        ("elementary", c_void_p),
        ("num_physical", c_uint64),
        ("physical", c_void_p),
    ]


class VtuArray(ctypes.Structure):
    """tg_vtu_array."""

    _fields_ = [
        ("name", c_char_p),
        ("type", c_int32),
        ("components", c_uint32),
        ("values", c_void_p),
        ("count", c_uint64),
    ]


class VtuBlock(ctypes.Structure):
    """tg_vtu_block."""

    _fields_ = [
        ("cell", c_int32),
        ("degree", c_uint32),
        ("nodes", c_void_p),
        ("len", c_uint64),
    ]


class GmshElementBlock(ctypes.Structure):
    """tg_gmsh_element_block."""

    _fields_ = [
        ("dim", c_int32),
        ("entity_tag", c_int32),
        ("element_type", c_int32),
        ("count", c_uint64),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with gridcl_amd.build.build_library() "
            "(the geometry path has no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().tg_last_error_message().decode(errors="replace")
