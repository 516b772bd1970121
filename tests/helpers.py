"""Shared test helpers: run the HIP path and the oracle on the same mesh and points."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from gridcl_amd import api
from gridcl_amd import meshes as M

LENS = {"points": lambda td, gd: gd, "jacobians": lambda td, gd: td * gd, "detj": lambda td, gd: 1,
        "pinv": lambda td, gd: td * gd, "facet_normals": lambda td, gd: gd,
        "normals": lambda td, gd: gd}


def gpu_factors(g: api.Geometry, b: int, ps: api.PointSet, factors, tile=16, options=0):
    """Cell-major [c][q][LEN] host arrays of the HIP evaluation."""
    import torch

    fields = api.geometric_factors_block(g, b, ps, factors, tile=tile, options=options)
    torch.cuda.synchronize()
    return {k: v.to_host_cell_major().astype(np.float64) for k, v in fields.items()}


def oracle_factors(g: api.Geometry, mesh: M.HostMesh, b: int, ps: api.PointSet, factors,
                   threads=0):
    """The oracle at the coordinates the geometry stores (widened to f64, plus the origin)."""
    blk = mesh.blocks[b]
    nodes = api.stored_nodes_f64(g)
    return O.eval_block(blk.cell, blk.degree, blk.family, mesh.gdim, nodes,
                        blk.cells.reshape(-1), ps.points, factors=factors, flags=blk.flags,
                        ref_normals=ps.facet_normals, threads=threads)


def rel_errs(got: dict, want: dict) -> dict:
    return {k: O.rel_err(got[k], want[k]) for k in got}


def pointwise_rel(got, want):
    """max over entries of |got - want| / |want| (for det J, whose size varies per cell)."""
    got = np.asarray(got).reshape(-1)
    want = np.asarray(want).reshape(-1)
    return float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300)))


def oracle_factors_sampled(g: api.Geometry, mesh: M.HostMesh, b: int, ps: api.PointSet, factors,
                           cells: np.ndarray, threads=0):
    """The oracle on the listed cells of block b only (cell-major [len(cells)][q][LEN])."""
    blk = mesh.blocks[b]
    nodes = api.stored_nodes_f64(g)
    rows = np.ascontiguousarray(blk.cells[cells]).reshape(-1)
    flags = None if blk.flags is None else np.ascontiguousarray(blk.flags[cells])
    return O.eval_block(blk.cell, blk.degree, blk.family, mesh.gdim, nodes, rows, ps.points,
                        factors=factors, flags=flags, ref_normals=ps.facet_normals,
                        threads=threads)
