"""Parity at the benchmark configurations' full sizes (needs the GPU).

Config 1: P1 tetrahedra, 32^3 x 6 = 196,608 cells, Keast-4 points, FP64 against the oracle.
Config 2: curved P2 tetrahedra, 64^3 x 6 = 1,572,864 cells, 14-point degree-5 rule: FP32 (the
per-cell shift and the FP64 fix-up) and FP64 against the oracle. Tolerances: FP64 1e-12, FP32
1e-6 for X, J and det J (block rel_err, reference.rs:776-817); the pointwise relative error of
det J is reported beside it.
"""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from tests.helpers import gpu_factors, oracle_factors, pointwise_rel, rel_errs

pytestmark = pytest.mark.gpu

NAMES = ["points", "jacobians", "detj"]


def test_config1_p1_tets_fp64(cuda):
    import torch

    mesh = M.config1_mesh()
    assert mesh.blocks[0].num_cells == 196_608 and mesh.num_nodes == 35_937
    g = api.geometry_from_host(mesh, dtype=torch.float64)
    ps = api.PointSet.tet_keast4()
    got = gpu_factors(g, 0, ps, NAMES)
    want = oracle_factors(g, mesh, 0, ps, NAMES)
    err = rel_errs(got, want)
    print("config 1 FP64 rel_err", err)
    for n, e in err.items():
        assert e <= 1e-12, (n, e)
    assert float(np.min(got["detj"])) > 0


@pytest.fixture(scope="module")
def config2():
    return M.config2_mesh()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_config2_p2_tets(cuda, config2, dtype):
    import torch

    mesh = config2
    assert mesh.blocks[0].num_cells == 1_572_864 and mesh.num_nodes == 2_146_689
    tdt = torch.float32 if dtype == "f32" else torch.float64
    g = api.geometry_from_host(mesh, dtype=tdt)
    ps = api.PointSet.tet_14()
    got = gpu_factors(g, 0, ps, NAMES)
    want = oracle_factors(g, mesh, 0, ps, NAMES)
    err = rel_errs(got, want)
    pw = pointwise_rel(got["detj"], want["detj"])
    print(f"config 2 {dtype} rel_err {err} det J pointwise {pw:.3e}")
    tol = 1e-6 if dtype == "f32" else 1e-12
    for n, e in err.items():
        assert e <= tol, (n, e)
    assert float(np.min(want["detj"])) > 0
