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


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_config3_q3_hexahedra(cuda, dtype):
    """Config 3: Q3 hexahedra 96^3 = 884,736 cells at 4^3 Gauss points (the sum-factorised
    kernel; FP32 with its last two sums on the matrix cores, and on the vector ALU under
    TG_OPT_NO_MFMA), against the oracle on every 211th cell. FP32 1e-6, FP64 1e-12 (X, J,
    det J)."""
    import torch

    from tests.helpers import oracle_factors_sampled

    mesh = M.config3_mesh()
    assert mesh.blocks[0].num_cells == 884_736 and mesh.num_nodes == 24_137_569
    tdt = torch.float32 if dtype == "f32" else torch.float64
    g = api.geometry_from_host(mesh, dtype=tdt)
    ps = api.PointSet.gauss(5, 7)
    assert len(ps) == 64
    cells = np.arange(0, mesh.blocks[0].num_cells, 211)
    want = oracle_factors_sampled(g, mesh, 0, ps, NAMES, cells)
    tol = 1e-6 if dtype == "f32" else 1e-12
    for path, opt in (("default", 0), ("vector ALU", api.OPT_NO_MFMA))[: 2 if dtype == "f32" else 1]:
        fields = api.geometric_factors_block(g, 0, ps, NAMES, options=opt)
        torch.cuda.synchronize()
        got = {}
        for n, f in fields.items():
            a = f.data.cpu().numpy().reshape(-1, 64, f.len, 16)  # [tile][q][LEN][16]
            got[n] = a[cells // 16, :, :, cells % 16].astype(np.float64)  # [cells][q][LEN]
        del fields
        err = rel_errs(got, want)
        pw = pointwise_rel(got["detj"], want["detj"])
        print(f"config 3 {dtype} {path} rel_err {err} det J pointwise {pw:.3e}")
        for n, e in err.items():
            assert e <= tol, (path, n, e)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_config4_hybrid(cuda, dtype):
    """Config 4: the mixed mesh (P2 tetrahedra, Q2 hexahedra, P2 prisms, P1 pyramids, P3 surface
    triangles with normals; 1.42 M cells, Gauss degree 4), each block's X, J, det J (and the
    triangles' normals) through the mesh plan's per-block launches and through its persistent
    launch, each eager and replayed from its graph (outputs poisoned with NaN before each run),
    against the oracle on every 97th cell of each block. FP32 1e-6 (block rel_err);
    FP64 5e-12: FP64 runs without a per-cell shift, and the P3 surface triangles' J sums ten
    products of coordinates about 40 cell widths from the origin with gradients up to ~10, so its
    rounding reaches ~1e-12 of the block's largest entry (1.03e-12 measured)."""
    import torch

    from tests.helpers import oracle_factors_sampled

    mesh = M.config4_mesh()
    tdt = torch.float32 if dtype == "f32" else torch.float64
    g = api.geometry_from_host(mesh, dtype=tdt)
    rule = api.QuadRule.for_grid(g, 4)
    factors = NAMES + ["normals"]
    tol = 1e-6 if dtype == "f32" else 5e-12
    worst = {}
    wants = {}
    for name, opt in (("per_block", api.OPT_PLAN_PER_BLOCK), ("persistent", api.OPT_PLAN_PERSISTENT)):
        plan = api.MeshPlan(g, rule, factors, options=opt)
        for mode in ("eager", "graph"):
            for fb in plan.fields:
                for f in fb.values():
                    f.data.fill_(float("nan"))
            print(f"config 4 {dtype} {name} {mode}", flush=True)
            if mode == "eager":
                plan.launch()
            else:
                # two replays with eager work between them; the second one is checked
                plan.graph_launch()
                torch.cuda.synchronize()
                for fb in plan.fields:
                    for f in fb.values():
                        f.data.fill_(float("nan"))
                plan.graph_launch()
            torch.cuda.synchronize()
            for b, blk in enumerate(mesh.blocks):
                ps = rule.get(blk.cell)
                cells = np.arange(0, blk.num_cells, 97)
                got = {}
                for n, f in plan.fields[b].items():
                    a = f.data.cpu().numpy().reshape(-1, len(ps), f.len, 16)
                    got[n] = a[cells // 16, :, :, cells % 16].astype(np.float64)
                if b not in wants:
                    wants[b] = oracle_factors_sampled(g, mesh, b, ps, list(got), cells)
                for n, e in rel_errs(got, wants[b]).items():
                    worst[(name, n)] = max(worst.get((name, n), 0.0), e)
                    assert e <= tol, (name, mode, b, blk.cell, n, e)
        del plan
    print(f"config 4 {dtype} rel_err {worst}")
