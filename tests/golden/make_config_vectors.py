"""Writes tests/golden/config_vectors.npz: X, J and det J of a few cells of each benchmark mesh,
evaluated by the FP64 oracle, with the cells' node coordinates and the mesh checksums.

    python tests/golden/make_config_vectors.py

Per configuration `c` (1, 2, 3, 5) the archive holds
  c{c}_cell, c{c}_degree    cell type and degree
  c{c}_points               the point set [nq][3]
  c{c}_cells                indices of the sampled cells in the block
  c{c}_nodes                their node coordinates [n][k][3] (FP64, as generated)
  c{c}_x, c{c}_j, c{c}_detj the oracle's X [n][nq][3], J [n][nq][9], det J [n][nq]
  c{c}_sum_nodes, c{c}_sum_dofmap  checksums of the whole generated mesh (sum of coordinates in
                            FP64, sum of the dofmap in integers)
The values are the oracle's on the exact FP64 coordinates; the GPU tests compare the FP64
kernels with them. Regenerate only after a deliberate change of the oracle or the generators.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gridcl_amd import meshes as M  # noqa: E402
from oracle import oracle as O  # noqa: E402

SAMPLES = 12


def tet14_points():
    """The 14-point degree-5 tetrahedron rule's points, from the library's C ABI (host call)."""
    from gridcl_amd import api

    return api.PointSet.tet_14().points.copy(), api.PointSet.tet_keast4().points.copy()


def gauss_hex_points(m: int) -> np.ndarray:
    g, _ = np.polynomial.legendre.leggauss(m)
    u = (g + 1.0) / 2.0
    return np.array([[u[i], u[j], u[k]] for k in range(m) for j in range(m) for i in range(m)])


def sample(mesh: M.HostMesh, points: np.ndarray, prefix: str, out: dict) -> None:
    blk = mesh.blocks[0]
    cells = np.unique(np.linspace(0, blk.num_cells - 1, SAMPLES).astype(np.int64))
    rows = blk.cells[cells]
    k = blk.nodes_per_cell
    nodes = mesh.nodes[rows.reshape(-1)].reshape(len(cells), k, 3)
    local = np.arange(len(cells) * k, dtype=np.uint32)
    res = O.eval_block(blk.cell, blk.degree, blk.family, 3, nodes.reshape(-1, 3), local, points,
                       factors=["points", "jacobians", "detj"])
    out[prefix + "cell"] = np.int64(blk.cell)
    out[prefix + "degree"] = np.int64(blk.degree)
    out[prefix + "points"] = points
    out[prefix + "cells"] = cells
    out[prefix + "nodes"] = nodes
    out[prefix + "x"] = res["points"]
    out[prefix + "j"] = res["jacobians"]
    out[prefix + "detj"] = res["detj"].reshape(len(cells), -1)
    out[prefix + "sum_nodes"] = np.float64(mesh.nodes.sum())
    out[prefix + "sum_dofmap"] = np.int64(blk.cells.astype(np.int64).sum())


def main() -> None:
    t14, keast = tet14_points()
    out: dict = {}
    sample(M.config1_mesh(), keast, "c1_", out)
    sample(M.config2_mesh(), t14, "c2_", out)
    sample(M.config3_mesh(), gauss_hex_points(4), "c3_", out)
    sample(M.config5_mesh(), t14, "c5_", out)
    np.savez_compressed(os.path.join(HERE, "config_vectors.npz"), **out)


if __name__ == "__main__":
    main()
