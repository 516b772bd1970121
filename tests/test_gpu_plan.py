"""The mixed-mesh plan: one persistent launch over every block of the hybrid mesh, and its
hipGraph replays, against the per-block launches and the oracle (needs the GPU).

The persistent kernel runs the register-resident per-block kernels' arithmetic, so its fields are
bitwise equal to tg_eval_block's with the sum-factorised hexahedron kernel switched off
(TG_OPT_NO_SUMFACT); graph replays are checked against the eager result after each replay, with
eager work and synchronisation in between."""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from tests.helpers import oracle_factors, rel_errs

pytestmark = pytest.mark.gpu
FACTORS = ["points", "jacobians", "detj", "normals"]


@pytest.fixture(scope="module")
def hybrid_small():
    return M.config4_mesh(seed=4, scale=0.01)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_plan_matches_per_block_and_oracle(cuda, hybrid_small, dtype):
    import torch

    h = hybrid_small
    tdt = torch.float32 if dtype == "f32" else torch.float64
    g = api.geometry_from_host(h, dtype=tdt)
    rule = api.QuadRule.for_grid(g, 4)
    plan = api.MeshPlan(g, rule, FACTORS, options=api.OPT_PLAN_PERSISTENT)
    # One workgroup per item at most; the small mesh has fewer items than resident workgroups.
    assert plan.persistent and 1 <= plan.grid <= plan.num_items
    plan.launch()
    torch.cuda.synchronize()
    tol = 1e-6 if dtype == "f32" else 1e-12
    for b, blk in enumerate(h.blocks):
        names = list(plan.fields[b])
        ps = rule.get(blk.cell)
        single = api.geometric_factors_block(g, b, ps, names, options=api.OPT_NO_SUMFACT)
        torch.cuda.synchronize()
        for n in names:
            assert np.array_equal(plan.fields[b][n].to_host_cell_major(),
                                  single[n].to_host_cell_major()), (b, n)
        got = {n: plan.fields[b][n].to_host_cell_major().astype(np.float64) for n in names}
        want = oracle_factors(g, h, b, ps, names)
        err = rel_errs(got, want)
        print(f"block {b} cell {blk.cell} p {blk.degree}: {err}")
        for n in ("points", "jacobians", "detj"):
            assert err[n] <= tol, (b, n, err[n])


def test_plan_graph_replays(cuda, hybrid_small):
    import torch

    g = api.geometry_from_host(hybrid_small, dtype=torch.float32)
    rule = api.QuadRule.for_grid(g, 4)
    plan = api.MeshPlan(g, rule, FACTORS, options=api.OPT_PLAN_PERSISTENT)
    plan.launch()
    torch.cuda.synchronize()
    eager = [{n: f.data.clone() for n, f in fb.items()} for fb in plan.fields]
    for rep in range(4):
        for fb in plan.fields:
            for f in fb.values():
                f.data.fill_(float("nan"))
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        for b, fb in enumerate(plan.fields):
            for n, f in fb.items():
                assert torch.equal(f.data, eager[b][n]), (rep, b, n)
        # Eager work between replays.
        plan.launch()
        torch.cuda.synchronize()


def test_plan_with_large_element_falls_back(cuda):
    """A block without a register-resident kernel (Q3 hexahedra, 64 nodes) makes the plan run
    one launch per block, with the same results."""
    import torch

    h = M.jitter(M.cube_box(5, 2, 3), 0.01, 3, interior_only=True)
    g = api.geometry_from_host(h, dtype=torch.float32)
    rule = api.QuadRule.for_grid(g, 4)
    plan = api.MeshPlan(g, rule, ["detj"], options=api.OPT_PLAN_PERSISTENT)
    assert not plan.persistent
    plan.launch()
    single = api.geometric_factors_block(g, 0, rule.get(5), ["detj"])
    torch.cuda.synchronize()
    assert np.array_equal(plan.fields[0]["detj"].to_host_cell_major(),
                          single["detj"].to_host_cell_major())


@pytest.mark.parametrize("serial", [False, True])
@pytest.mark.parametrize("graph", [False, True])
def test_plan_per_block_option_matches_eval(cuda, hybrid_small, serial, graph):
    """TG_OPT_PLAN_PER_BLOCK: one launch per block with tg_eval_block's kernels, forked onto
    streams of their own and joined back (or, with TG_OPT_PLAN_SERIAL, in order on the caller's
    stream), eagerly or replayed from a graph (three replays, outputs cleared before each);
    equal to per-block evaluation bitwise."""
    import torch

    g = api.geometry_from_host(hybrid_small, dtype=torch.float32)
    rule = api.QuadRule.for_grid(g, 4)
    opts = api.OPT_PLAN_PER_BLOCK | (api.OPT_PLAN_SERIAL if serial else 0)
    plan = api.MeshPlan(g, rule, FACTORS, options=opts)
    assert not plan.persistent
    for _ in range(3):
        for fb in plan.fields:
            for f in fb.values():
                f.data.zero_()
        plan.graph_launch() if graph else plan.launch()
        torch.cuda.synchronize()
        for b, blk in enumerate(hybrid_small.blocks):
            names = list(plan.fields[b])
            single = api.geometric_factors_block(g, b, rule.get(blk.cell), names)
            torch.cuda.synchronize()
            for n in names:
                assert np.array_equal(plan.fields[b][n].to_host_cell_major(),
                                      single[n].to_host_cell_major()), (b, n)
