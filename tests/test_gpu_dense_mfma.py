"""P3 tetrahedra (20 nodes) in FP32 on the dense matrix-core kernel (dense_mfma.hip), on the GPU.

The kernel takes every call for J on FP32 P3 tetrahedra with tiles of 16 and at most 128
points. Each v_mfma_f32_16x16x4f32 step adds its four products to the accumulator in node order,
each rounded once, so the sums over the 20 nodes are the register kernel's chain of fused
multiply-adds from the first node: the values equal the register kernel's (TG_OPT_NO_MFMA) bit
for bit, and both are held to the oracle with the tolerance for 20-node FP32 sums (4e-6, block
rel_err, tests/test_gpu_parity.py). Calls that fall back (more than 128 points, X or det J
alone, forced strategies) run the other kernels."""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from tests.helpers import gpu_factors, oracle_factors, pointwise_rel, rel_errs
from tests.test_gpu_parity import volume_mesh

pytestmark = pytest.mark.gpu
TET = M.TETRAHEDRON
NAMES = ["points", "jacobians", "detj"]


def _sets():
    # Keast-4 (4 points), the 14-point rule, Gauss degrees 4, 6 and 8 (27, 64, 125 points: one
    # to eight blocks of 16 points, the last one partial).
    return [("keast4", api.PointSet.tet_keast4()), ("tet14", api.PointSet.tet_14()),
            ("gauss4", api.PointSet.gauss(TET, 4)), ("gauss6", api.PointSet.gauss(TET, 6)),
            ("gauss8", api.PointSet.gauss(TET, 8))]


@pytest.mark.parametrize("idx", range(5))
def test_against_oracle_and_register_kernel(cuda, idx):
    import torch

    name, ps = _sets()[idx]
    mesh = volume_mesh(TET, 3)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    got = gpu_factors(g, 0, ps, NAMES)
    reg = gpu_factors(g, 0, ps, NAMES, options=api.OPT_NO_MFMA)
    want = oracle_factors(g, mesh, 0, ps, NAMES)
    err, err_reg, diff = rel_errs(got, want), rel_errs(reg, want), rel_errs(got, reg)
    print(f"P3 tet {name} ({len(ps)} points): matrix cores {err}, register kernel {err_reg}, "
          f"between {diff}")
    for n in NAMES:
        assert err[n] <= 4e-6, (n, err[n])
        assert np.array_equal(got[n], reg[n]), n


@pytest.mark.parametrize("cells", [1, 17, 50])
def test_ragged_blocks(cuda, cells):
    """Blocks whose cell count is not a multiple of the 16-cell tile (and of the four tiles of a
    workgroup): padding lanes repeat the last cell; every real cell against the oracle."""
    import torch

    mesh = volume_mesh(TET, 3)
    b = mesh.blocks[0]
    mesh.blocks[0] = M.HostBlock(b.cell, b.degree, b.cells[:cells].copy())
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    ps = api.PointSet.gauss(TET, 6)
    got = gpu_factors(g, 0, ps, NAMES)
    want = oracle_factors(g, mesh, 0, ps, NAMES)
    for n, e in rel_errs(got, want).items():
        assert e <= 4e-6, (cells, n, e)


def test_factor_subsets_and_fallbacks(cuda):
    """Every subset of X, J and det J runs on the matrix cores and equals the full call bit for
    bit, and X alone and det J alone (the measure's call) equal the register kernel's
    (TG_OPT_NO_MFMA); more than 128 points (Gauss degree 10, 216) and TG_OPT_COOPERATIVE /
    TG_OPT_FORCE_GENERIC keep their kernels."""
    import torch

    mesh = volume_mesh(TET, 3)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    ps = api.PointSet.gauss(TET, 6)
    full = gpu_factors(g, 0, ps, NAMES)
    for subset in (["jacobians"], ["points", "jacobians"], ["jacobians", "detj"], ["points"],
                   ["detj"], ["points", "detj"]):
        part = gpu_factors(g, 0, ps, subset)
        for n in subset:
            assert np.array_equal(part[n], full[n]), (subset, n)
    reg = gpu_factors(g, 0, ps, NAMES, options=api.OPT_NO_MFMA)
    for n in ("points", "detj"):
        alone = gpu_factors(g, 0, ps, [n])[n]
        assert np.array_equal(alone, reg[n]), n
    big = api.PointSet.gauss(TET, 10)
    assert len(big) > 128
    assert all(np.array_equal(a, b) for a, b in zip(
        gpu_factors(g, 0, big, NAMES).values(),
        gpu_factors(g, 0, big, NAMES, options=api.OPT_NO_MFMA).values()))
    for opt in (api.OPT_COOPERATIVE, api.OPT_FORCE_GENERIC):
        forced = gpu_factors(g, 0, ps, NAMES, options=opt)
        for n, e in rel_errs(forced, oracle_factors(g, mesh, 0, ps, NAMES)).items():
            assert e <= 1e-5, (opt, n, e)


def test_fp64_fixup_on_near_degenerate_cells(cuda):
    """Near-degenerate points (|det J| < tau |J|_F^3) get det J in FP64 on the matrix-core path
    too: on a box flattened in z by 1e-4 the pointwise det J matches the oracle to 1e-6 where the
    FP32 value alone does not."""
    import torch

    mesh = volume_mesh(TET, 3, n=2)
    x = mesh.nodes.copy()
    x[:, 2] = 0.3 + 1e-4 * x[:, 2]
    mesh.nodes = x
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    ps = api.PointSet.tet_14()
    want = oracle_factors(g, mesh, 0, ps, ["detj"])["detj"]
    g.set_fixup_threshold(1e-2)
    fixed = gpu_factors(g, 0, ps, NAMES)["detj"]
    plain = gpu_factors(g, 0, ps, NAMES, options=api.OPT_NO_FP64_FIXUP)["detj"]
    e_fixed, e_plain = pointwise_rel(fixed, want), pointwise_rel(plain, want)
    print("P3 near-degenerate det J pointwise rel err: fixed", e_fixed, "fp32 only", e_plain)
    assert e_fixed <= 1e-6
    assert e_plain > e_fixed


def test_plans_keep_the_block_on_its_kernel(cuda):
    """A mesh plan over P3 tetrahedra runs the block's own launch (the persistent kernel is not
    taken even when asked for), eagerly and from its graph, equal to tg_eval_block bit for bit."""
    import torch

    mesh = volume_mesh(TET, 3)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    rule = api.QuadRule.gauss([TET], 6)
    single = gpu_factors(g, 0, rule.get(TET), NAMES)
    for opt in (api.OPT_PLAN_PERSISTENT, api.OPT_PLAN_PER_BLOCK):
        plan = api.MeshPlan(g, rule, NAMES, options=opt)
        assert not plan.persistent
        for launch in (plan.launch, plan.graph_launch):
            for f in plan.fields[0].values():
                f.data.fill_(float("nan"))
            launch()
            torch.cuda.synchronize()
            for n in NAMES:
                assert np.array_equal(plan.fields[0][n].to_host_cell_major(), single[n]), (opt, n)


def test_full_size_mesh(cuda):
    """The bench leg's mesh, curved P3 tetrahedra 48^3 x 6 (663,552 cells, 3,048,625 nodes), at
    the 14-point rule and Gauss degree 6: the matrix-core fields equal the register kernel's bit
    for bit over the whole mesh (compared on the device), and every 97th cell matches the oracle
    (4e-6, block rel_err, as for 20-node FP32 sums above)."""
    import torch

    from tests.helpers import oracle_factors_sampled

    mesh = M.p3tet_mesh()
    C = mesh.blocks[0].num_cells
    assert C == 663_552 and mesh.num_nodes == 3_048_625
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    cells = np.arange(0, C, 97)
    for name, ps in (("tet14", api.PointSet.tet_14()), ("gauss6", api.PointSet.gauss(TET, 6))):
        nq = len(ps)
        fields = api.geometric_factors_block(g, 0, ps, NAMES)
        reg = api.geometric_factors_block(g, 0, ps, NAMES, options=api.OPT_NO_MFMA)
        torch.cuda.synchronize()
        for n in NAMES:
            assert torch.equal(fields[n].data, reg[n].data), (name, n)
        del reg
        idx = torch.from_numpy(cells).to(fields["points"].data.device)
        got = {}
        for n, f in fields.items():
            a = f.data.view(-1, nq, f.len, 16)  # [tile][q][LEN][16]
            got[n] = a[idx // 16, :, :, idx % 16].cpu().numpy().astype(np.float64)
        del fields
        want = oracle_factors_sampled(g, mesh, 0, ps, NAMES, cells)
        err = rel_errs(got, want)
        print(f"P3 tets 48^3 x 6 {name}: rel_err against the oracle {err}")
        for n, e in err.items():
            assert e <= 4e-6, (name, n, e)
        assert float(np.min(want["detj"])) > 0


@pytest.mark.parametrize("idx", [1, 3])
def test_fp64_stored_nodes(cuda, idx):
    """FP32 P3 tetrahedra on FP64-stored nodes take the matrix-core kernel too (each offset from
    node 0 taken in FP64 and rounded once to FP32): equal bit for bit to the register kernel on
    the same nodes (TG_OPT_NO_MFMA), within 4e-6 of the exact mesh (the oracle on the FP64 nodes;
    the tolerance of the 20-term FP32 sums on FP64 nodes, tests/test_gpu_nodes64.py), and not the
    values of the FP32-node geometry."""
    import torch

    from tests.test_gpu_nodes64 import exact_oracle

    name, ps = _sets()[idx]
    mesh = volume_mesh(TET, 3)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    got = gpu_factors(g, 0, ps, NAMES)
    reg = gpu_factors(g, 0, ps, NAMES, options=api.OPT_NO_MFMA)
    for n in NAMES:
        assert np.array_equal(got[n], reg[n]), n
    err = rel_errs(got, exact_oracle(mesh, ps, NAMES))
    print(f"P3 tet {name} on FP64 nodes: {err}")
    for n in NAMES:
        assert err[n] <= 4e-6, (n, err[n])
    f32 = gpu_factors(api.geometry_from_host(mesh, dtype=torch.float32), 0, ps, NAMES)
    assert not np.array_equal(got["jacobians"], f32["jacobians"])


PYR, PRISM, HEX = M.PYRAMID, M.PRISM, M.HEXAHEDRON


def _scattered(cell, n, seed=3):
    """n seeded points in the unit cube (no tensor structure, so the sum-factorised kernels do
    not take them), equal weights."""
    pts = M.splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    return api.PointSet.custom(cell, pts, np.full(n, 1.0 / n))


# (label, cell, degree, point set, whether a register-resident kernel exists)
SHAPES = [("P3 tet", TET, 3, lambda: api.PointSet.gauss(TET, 4), True),
          ("Q2 hex", HEX, 2, lambda: _scattered(HEX, 40), True),
          ("P3 pyramid", PYR, 3, lambda: api.PointSet.gauss(PYR, 4), False),
          ("P4 tet", TET, 4, lambda: api.PointSet.gauss(TET, 5), False),
          ("P3 prism", PRISM, 3, lambda: api.PointSet.gauss(PRISM, 5), False),
          ("P5 tet", TET, 5, lambda: api.PointSet.gauss(TET, 6), False),
          ("Q3 hex", HEX, 3, lambda: _scattered(HEX, 100), True),
          ("P4 prism", PRISM, 4, lambda: api.PointSet.gauss(PRISM, 5), False),
          ("P6 tet", TET, 6, lambda: api.PointSet.gauss(TET, 7), False),
          # the runtime-k form: more than 84 nodes, or a count the register-operand form lacks
          ("P4 pyramid", PYR, 4, lambda: api.PointSet.gauss(PYR, 5), False),
          ("P7 tet", TET, 7, lambda: api.PointSet.gauss(TET, 8), False),
          ("Q4 hex", HEX, 4, lambda: _scattered(HEX, 64), False),
          ("P5 prism", PRISM, 5, lambda: api.PointSet.gauss(PRISM, 6), False),
          ("P8 tet", TET, 8, lambda: api.PointSet.gauss(TET, 9), False)]


@pytest.mark.parametrize("idx", range(len(SHAPES)))
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_every_shape_against_oracle_and_node_order_kernels(cuda, idx, dtype):
    """Every cell the kernel takes (20 to 165 nodes; the runtime-k form from P4 pyramids and
    above 84 nodes): X, J and det J equal the runtime-K kernel's (TG_OPT_FORCE_GENERIC: node
    order from node 0) bit for bit, and the register kernel's where the cell has one
    (TG_OPT_NO_MFMA); against the oracle within 4e-6 where a register kernel exists (the dense
    FP32 sums above), otherwise within the reference's scaled tolerances on its f32 base 1e-5
    (tests/test_gpu_coop.py; the base itself up to the degree-4 element). FP64 geometries take
    the FP64 matrix cores (v_mfma_f64_16x16x4f64) from 20 nodes, with the same checks against
    the FP64 kernels (node order from node 0, no shift) and the FP64 base 1e-12. X alone and det J
    alone equal the full call."""
    import torch

    from tests.test_gpu_coop import scaled_tolerances

    label, cell, p, make, has_reg = SHAPES[idx]
    ps = make()
    assert len(ps) <= 128
    mesh = volume_mesh(cell, p, n=2)
    f64 = dtype == "f64"
    g = api.geometry_from_host(mesh, dtype=torch.float64 if f64 else torch.float32)
    got = gpu_factors(g, 0, ps, NAMES)
    gen = gpu_factors(g, 0, ps, NAMES, options=api.OPT_FORCE_GENERIC)
    for n in NAMES:
        assert np.array_equal(got[n], gen[n]), (label, n)
    if has_reg:
        reg = gpu_factors(g, 0, ps, NAMES, options=api.OPT_NO_MFMA)
        for n in NAMES:
            assert np.array_equal(got[n], reg[n]), (label, n)
    want = oracle_factors(g, mesh, 0, ps, NAMES)
    err = rel_errs(got, want)
    base = 1e-12 if f64 else 1e-5
    if has_reg and not f64:
        tol = {n: 4e-6 for n in NAMES}
    elif f64 and (cell, p) == (PYR, 4):
        # The library's and the oracle's pyramid tables both come from a Vandermonde solve whose
        # conditioning costs a digit at degree 4 (tests/test_refcell.py allows 1e-11 for the
        # values and 1e-10 for the derivatives there); the sums' own rounding is 8e-15 of max |J|
        # on this mesh. Measured: J 6.2e-12, det J 2.2e-11.
        tol = {"points": 1e-11, "jacobians": 1e-10, "detj": 1e-10}
    else:
        tol = scaled_tolerances(cell, p, ps.points, base, want["jacobians"])
    print(f"{label} {dtype} ({mesh.blocks[0].nodes_per_cell} nodes, {len(ps)} points): {err}")
    for n in NAMES:
        assert err[n] <= tol[n], (label, n, err[n], tol[n])
    for n in ("points", "detj"):
        assert np.array_equal(gpu_factors(g, 0, ps, [n])[n], got[n]), (label, n)


@pytest.mark.parametrize("idx", range(len(SHAPES)))
def test_every_shape_on_fp64_stored_nodes(cuda, idx):
    """FP32 geometries on FP64-stored nodes, every cell the kernel takes: each offset from node 0
    taken in FP64 and rounded once, so X, J and det J equal the runtime-K kernel's on the same
    nodes (TG_OPT_FORCE_GENERIC) bit for bit; X alone and det J alone (the forms whose table
    loads leave out the other kind) equal the full call; and the values are not those of the
    FP32-node geometry (the FP64 nodes were read)."""
    import torch

    label, cell, p, make, _ = SHAPES[idx]
    ps = make()
    mesh = volume_mesh(cell, p, n=2)
    g = api.geometry_from_host(mesh, dtype=torch.float32, nodes_f64=True)
    got = gpu_factors(g, 0, ps, NAMES)
    gen = gpu_factors(g, 0, ps, NAMES, options=api.OPT_FORCE_GENERIC)
    for n in NAMES:
        assert np.array_equal(got[n], gen[n]), (label, n)
    for n in ("points", "detj"):
        assert np.array_equal(gpu_factors(g, 0, ps, [n])[n], got[n]), (label, n)
    f32 = gpu_factors(api.geometry_from_host(mesh, dtype=torch.float32), 0, ps, NAMES)
    assert not np.array_equal(got["jacobians"], f32["jacobians"]), label


@pytest.mark.parametrize("idx", range(len(SHAPES)))
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_pinv_and_facet_normals_on_every_shape(cuda, idx, dtype):
    """K with X, J and det J, K alone, and facet normals at the facet Gauss points, on the matrix
    cores:
    bit for bit the runtime-K kernel's (and the register kernel's where there is one), and
    against the oracle within the reference's scaled tolerances (K with kappa(J),
    tests/test_gpu_coop.py; 1e-12 / 1e-5 base for FP64 / FP32; the degree-4 pyramid's tables as
    above), the facet normals with K's tolerance (1.3e-5 measured for FP32 P8 tetrahedra, whose
    K tolerance is 3e-5 there)."""
    import torch

    from tests.test_gpu_coop import scaled_tolerances

    label, cell, p, make, has_reg = SHAPES[idx]
    f64 = dtype == "f64"
    mesh = volume_mesh(cell, p, n=2)
    g = api.geometry_from_host(mesh, dtype=torch.float64 if f64 else torch.float32)
    names = NAMES + ["pinv"]
    fps = api.PointSet.facet_gauss(cell, 2)
    for ps, fac in ((make(), names), (make(), ["pinv"]), (fps, ["facet_normals"])):
        got = gpu_factors(g, 0, ps, fac)
        gen = gpu_factors(g, 0, ps, fac, options=api.OPT_FORCE_GENERIC)
        for n in fac:
            assert np.array_equal(got[n], gen[n]), (label, n)
        if has_reg:
            reg = gpu_factors(g, 0, ps, fac, options=api.OPT_NO_MFMA)
            for n in fac:
                assert np.array_equal(got[n], reg[n]), (label, n)
        want = oracle_factors(g, mesh, 0, ps, fac)
        err = rel_errs(got, want)
        base = 1e-12 if f64 else 1e-5
        jw = oracle_factors(g, mesh, 0, ps, ["jacobians"])["jacobians"]
        tol = scaled_tolerances(cell, p, ps.points, base, jw)
        tol["facet_normals"] = tol["pinv"]  # K^T n / |K^T n|: K's tolerance
        if f64 and (cell, p) == (PYR, 4):  # the tables' digit, as in the test above
            tol = {n: 1e-9 if n in ("pinv", "facet_normals") else 1e-10 for n in tol}
        print(f"{label} {dtype} {fac}: {err}")
        for n in fac:
            assert err[n] <= tol[n], (label, n, err[n], tol[n])


@pytest.mark.parametrize("cell,p,deg", [(TET, 3, 6), (TET, 7, 8), (PRISM, 5, 6)])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_ragged_and_clamped_blocks(cuda, cell, p, deg, dtype):
    """The register-operand form (P3 tetrahedra) and the runtime-k form (P7 tetrahedra, P5 prisms;
    FP64 always) on blocks of 1, 17 and 37 cells (the last tile, and the last (tile, block of
    points) item, part-full) with dofmap entries past the last node (read as the last node,
    evector.rs:187-191): bit for bit the runtime-K kernel, which clamps the same way."""
    import torch

    mesh = volume_mesh(cell, p, n=2)
    b = mesh.blocks[0]
    cells = b.cells.copy()
    cells[3, 5] = 10_000_000
    cells[min(20, len(cells) - 1), 0] = 0xFFFFFFFF
    ps = api.PointSet.gauss(cell, deg)
    assert len(ps) <= 128
    tdt = torch.float64 if dtype == "f64" else torch.float32
    for n in (1, 17, 37):
        m = M.HostMesh(mesh.gdim, mesh.nodes, [M.HostBlock(b.cell, b.degree, cells[:n])])
        g = api.geometry_from_host(m, dtype=tdt)
        names = NAMES + ["pinv"]
        got = gpu_factors(g, 0, ps, names)
        gen = gpu_factors(g, 0, ps, names, options=api.OPT_FORCE_GENERIC)
        for nm in names:
            assert np.array_equal(got[nm], gen[nm]), (n, nm)
