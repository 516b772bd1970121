"""FP32 arithmetic on FP64-stored nodes (tg_geometry_set_nodes_f64), on the GPU.

The reference's FP32 path reads FP32 nodes, and its parity convention compares with the FP64
evaluation of those rounded nodes (tetraloma-testing/src/reference.rs:47-69). Here the reference
values are the exact FP64 mesh: with the nodes stored in FP64 each cell's offsets carry one FP32
rounding of their own size, so what is left is the FP32 contraction's rounding. Errors are the
block metric rel_err = max|got - want| / max|want| (reference.rs:776-817).

Tolerances: 1e-6 for X, J and det J where the FP32 sums have at most 18 terms or the sums are
taken in FP64 (the sum-factorised hexahedra), 4e-6 for the dense 20- and 27-term FP32 sums (P3
tetrahedra, Q2 hexahedra outside the sum-factorised point sets), 1e-5 for K and normals (the
reference's FP32 tolerance). On the full configuration meshes the target is the accuracy goal of
the path: X, J and det J within 1e-6 of the FP64 kernel on the same FP64 mesh.
"""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from oracle import oracle as O
from tests.helpers import gpu_factors, rel_errs
from tests.test_gpu_parity import (HEX, INTERVAL, PRISM, PYR, QUAD, TET, TRI, factor_names,
                                   make_case)

pytestmark = pytest.mark.gpu

CASES = ([("vol", TET, p) for p in (1, 2, 3)] + [("vol", HEX, p) for p in (1, 2, 3)]
         + [("vol", PRISM, p) for p in (1, 2)] + [("vol", PYR, p) for p in (1, 2)]
         + [("surf3", TRI, p) for p in (1, 2, 3)] + [("surf2", QUAD, p) for p in (1, 2)]
         + [("line3", INTERVAL, p) for p in (2, 3)])


def exact_oracle(mesh, ps, names, b=0):
    blk = mesh.blocks[b]
    return O.eval_block(blk.cell, blk.degree, blk.family, mesh.gdim, mesh.nodes,
                        blk.cells.reshape(-1), ps.points, factors=names, flags=blk.flags,
                        ref_normals=ps.facet_normals)


@pytest.mark.parametrize("kind,cell,p", CASES)
def test_factors_against_the_exact_mesh(cuda, kind, cell, p):
    import torch

    mesh = make_case(kind, cell, p)
    td, gd = M.TDIM[cell], mesh.gdim
    ps = api.PointSet.gauss(cell, 2 * p)
    names = factor_names(td, gd)
    want = exact_oracle(mesh, ps, names)
    g64n = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    g32 = api.geometry_from_host(mesh, dtype=torch.float32)
    err = rel_errs(gpu_factors(g64n, 0, ps, names), want)
    err32 = rel_errs(gpu_factors(g32, 0, ps, names), want)
    print(f"{kind} cell={cell} p={p} FP64 nodes {err} FP32 nodes {err32}")
    k = mesh.blocks[0].nodes_per_cell
    # K among the factors keeps the hexahedra on the dense kernels (FP32 sums over all k nodes)
    base = 1e-6 if k <= 18 else 4e-6
    for name in ("points", "jacobians", "detj"):
        assert err[name] <= base, (name, err[name])
    for name in ("pinv", "normals"):
        if name in err:
            assert err[name] <= 1e-5, (name, err[name])


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("m", [2, 3, 4])
def test_sum_factorised_hexahedra_round_the_fp64_kernel(cuda, p, m):
    """The vector-ALU sum-factorised hexahedra on FP64 nodes (TG_OPT_NO_MFMA) take their sums in
    FP64: X, J and det J are the FP64 kernel's values rounded to FP32, bit for bit (the box's
    origin, 0.5, is exact in FP32)."""
    import torch

    mesh = make_case("vol", HEX, p)
    ps = api.PointSet.gauss(HEX, 2 * m - 1)
    names = ["points", "jacobians", "detj"]
    g64 = api.geometry_from_host(mesh, dtype=torch.float64)
    assert all(float(o) == float(np.float32(o)) for o in g64.origin)
    want = gpu_factors(g64, 0, ps, names, options=api.OPT_SUMFACT_ALWAYS)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    got = gpu_factors(g, 0, ps, names, options=api.OPT_SUMFACT_ALWAYS | api.OPT_NO_MFMA)
    for n in names:
        assert np.array_equal(got[n], want[n].astype(np.float32).astype(np.float64)), n


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("m", [2, 3, 4])
def test_matrix_core_hexahedra_on_fp64_nodes(cuda, p, m):
    """The matrix-core kernel on FP64 nodes (first sums in FP64, their hi and lo FP32 parts
    through the products) against the FP64 kernel on the same mesh: X, J and det J within 1e-6;
    and the FP32 arithmetic is visible, so this is not the vector-ALU FP64 path."""
    import torch

    mesh = make_case("vol", HEX, p)
    ps = api.PointSet.gauss(HEX, 2 * m - 1)
    names = ["points", "jacobians", "detj"]
    g64 = api.geometry_from_host(mesh, dtype=torch.float64)
    want = gpu_factors(g64, 0, ps, names, options=api.OPT_SUMFACT_ALWAYS)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    got = gpu_factors(g, 0, ps, names, options=api.OPT_SUMFACT_ALWAYS)
    valu = gpu_factors(g, 0, ps, names, options=api.OPT_SUMFACT_ALWAYS | api.OPT_NO_MFMA)
    err = rel_errs(got, want)
    print(f"p={p} m={m} matrix cores on FP64 nodes {err}")
    for n in names:
        assert err[n] <= 1e-6, (n, err[n])
    assert not np.array_equal(got["jacobians"], valu["jacobians"])


@pytest.mark.parametrize("p,m", [(2, 3), (3, 4)])
def test_matrix_core_factor_subsets_on_fp64_nodes(cuda, p, m):
    """The matrix-core kernel on FP64 nodes asked for one factor at a time (det J alone, as the
    measure asks; X alone, as the equispaced conversion does; J alone) gives the values of the
    call for all three, bit for bit; and the measure (det J alone through the same kernel) is
    the FP64 geometry's within 1e-6."""
    import torch

    mesh = make_case("vol", HEX, p)
    ps = api.PointSet.gauss(HEX, 2 * m - 1)
    names = ["points", "jacobians", "detj"]
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    full = gpu_factors(g, 0, ps, names)
    for n in names:
        one = gpu_factors(g, 0, ps, [n])
        assert np.array_equal(one[n], full[n]), n
    vol = api.measure_block(g, 0, ps)
    vol64 = api.measure_block(api.geometry_from_host(mesh, dtype=torch.float64), 0, ps)
    assert abs(vol - vol64) <= 1e-6 * abs(vol64), (vol, vol64)


@pytest.mark.parametrize("cell,p", [(TET, 2), (HEX, 2), (PRISM, 2)])
def test_generic_and_dense_kernels(cuda, cell, p):
    """The runtime-K kernel (forced) and the dense hexahedron kernel (sum factorisation off) on
    FP64 nodes, against the exact mesh."""
    import torch

    mesh = make_case("vol", cell, p)
    ps = api.PointSet.gauss(cell, 2 * p)
    names = ["points", "jacobians", "detj"]
    want = exact_oracle(mesh, ps, names)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    for opt in (api.OPT_FORCE_GENERIC, api.OPT_NO_SUMFACT):
        err = rel_errs(gpu_factors(g, 0, ps, names, options=opt), want)
        print(f"cell={cell} p={p} options={opt} {err}")
        tol = 1e-6 if mesh.blocks[0].nodes_per_cell <= 18 else 4e-6
        for name, e in err.items():
            assert e <= tol, (opt, name, e)


def test_checks_and_detach(cuda):
    """FP64 geometries refuse FP64 nodes; the shape must match; detaching restores the FP32-node
    results bit for bit."""
    import torch

    mesh = make_case("vol", TET, 2)
    ps = api.PointSet.tet_14()
    g64 = api.geometry_from_host(mesh, dtype=torch.float64)
    with pytest.raises(api.UnsupportedPrecision):
        g64.set_nodes_f64(g64.nodes)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    with pytest.raises(api.LengthMismatch):
        g.set_nodes_f64(g.nodes_f64[:-1])
    wide = gpu_factors(g, 0, ps, ["jacobians"])["jacobians"]
    g.set_nodes_f64(None)
    plain = gpu_factors(g, 0, ps, ["jacobians"])["jacobians"]
    ref = gpu_factors(api.geometry_from_host(mesh, dtype=torch.float32), 0, ps,
                      ["jacobians"])["jacobians"]
    assert np.array_equal(plain, ref)
    assert not np.array_equal(wide, plain)


def test_measure_and_plans(cuda):
    """The measure (its det J from the evaluation kernels) and the plans use the FP64 nodes: the
    curved unit box measures 1, and a plan's fields equal tg_eval_block's bit for bit; the
    persistent kernel, which reads the FP32 node array, is not taken."""
    import torch

    mesh = M.config4_mesh(seed=4, scale=0.002)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    rule = api.QuadRule.for_grid(g, 4)
    plan = api.MeshPlan(g, rule, ["points", "jacobians", "detj"], options=api.OPT_PLAN_PERSISTENT)
    assert not plan.persistent
    plan.launch()
    torch.cuda.synchronize()
    for b, blk in enumerate(mesh.blocks):
        single = api.geometric_factors_block(g, b, rule.get(blk.cell),
                                             list(plan.fields[b].keys()))
        torch.cuda.synchronize()
        for name, f in plan.fields[b].items():
            assert torch.equal(f.data, single[name].data), (b, name)
    box = make_case("vol", TET, 2)
    gb = api.geometry_from_host(box, dtype=torch.float32, nodes_f64=True)
    vol = api.measure_block(gb, 0, api.PointSet.tet_14())
    assert abs(vol - 1.0) <= 1e-6, vol


def _device_rel(a, b):
    """rel_err of two fields of the same layout, on the device in f64."""
    import torch

    d = (a.data.to(torch.float64) - b.data).abs().max()
    return float(d / b.data.abs().max())


@pytest.mark.parametrize("config", [2, 3])
def test_configurations_against_the_fp64_kernel(cuda, config):
    """Config 2 (P2 tetrahedra 64^3 x 6, 14 points) and config 3 (Q3 hexahedra 96^3, 4^3 points)
    at full size: X, J and det J of the FP32 kernels on FP64 nodes against the FP64 kernel on the
    same FP64 mesh, over every cell and point (the FP32 node array alone gives J ~1e-5 on config
    2). Limit 1e-6."""
    import torch

    if config == 2:
        mesh, ps = M.config2_mesh(), api.PointSet.tet_14()
    else:
        mesh, ps = M.config3_mesh(), api.PointSet.gauss(HEX, 7)
    names = ["points", "jacobians", "detj"]
    g64 = api.geometry_from_host(mesh, dtype=torch.float64)
    want = api.geometric_factors_block(g64, 0, ps, names)
    del g64
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    got = api.geometric_factors_block(g, 0, ps, names)
    torch.cuda.synchronize()
    err = {n: _device_rel(got[n], want[n]) for n in names}
    g.set_nodes_f64(None)
    plain = api.geometric_factors_block(g, 0, ps, names)
    torch.cuda.synchronize()
    err32 = {n: _device_rel(plain[n], want[n]) for n in names}
    print(f"config {config}: FP32 on FP64 nodes {err}; FP32 nodes {err32}")
    for n, e in err.items():
        assert e <= 1e-6, (n, e)
