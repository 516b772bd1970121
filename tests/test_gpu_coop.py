"""The cooperative strategy (tetraloma-geometry/src/cooperative/factors.rs:100-183) on the GPU:
each cell's node sums split among L lanes (L the largest power of two not above min(32, k)),
leaf l summing nodes l, l + L, ... from the product of node l, the leaves added by a pairwise
tree over cross-lane shuffles. Against the FP64 oracle (errors are the block metric rel_err of
tetraloma-testing/src/reference.rs:776-817):

- forced (TG_OPT_COOPERATIVE) on every element of the parity suite, 1 to 27 nodes per cell and
  the 64-node Q3 hexahedron: FP64 1e-12, FP32 the reference's f32 base tolerance 1e-5;
- by default on the tetrahedra of degree 5-8 (FP32) and 6-8 (FP64), up to 165 nodes per cell
  (at least 56 and 84 nodes, where it measured faster than the runtime-K kernel; the
  reference's GPU policy takes it from degree 7 and 6, degree-policy.toml:77,86), and the
  runtime-K kernel on the same meshes for comparison. These degrees take the reference's scaled
  tolerances (tetraloma-testing/src/reference.rs:17-22, 478-560): the base tolerance times
  max(1, A / A_ref) with A the element's amplification factor at the test's points (the
  Lebesgue function for X; max_d sum_j |d_d psi_j| for J and what derives from it) and A_ref the
  same of the degree-4 element, times max(1, n / n_ref) nodes for J, det J and K, and for K
  also the condition number kappa_F(J) / tdim.
"""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from oracle import oracle as O
from tests.helpers import gpu_factors, oracle_factors, rel_errs
from tests.test_gpu_parity import CASES, TET, factor_names, make_case, volume_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,cell,p", CASES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_cooperative_matches_oracle(cuda, kind, cell, p, dtype):
    import torch

    mesh = make_case(kind, cell, p)
    td, gd = M.TDIM[cell], mesh.gdim
    tdt = torch.float64 if dtype == "f64" else torch.float32
    g = api.geometry_from_host(mesh, dtype=tdt)
    ps = api.PointSet.gauss(cell, 2 * p)
    names = factor_names(td, gd)
    got = gpu_factors(g, 0, ps, names, options=api.OPT_COOPERATIVE)
    want = oracle_factors(g, mesh, 0, ps, names)
    err = rel_errs(got, want)
    print(f"{kind} cell={cell} p={p} {dtype} k={mesh.blocks[0].nodes_per_cell} rel_err={err}")
    tol = 1e-12 if dtype == "f64" else 1e-5
    for name, e in err.items():
        assert e <= tol, (name, e)


def amplification(cell, degree, pts):
    """(Lebesgue function, max_d sum_j |d_d psi_j|), each the maximum over the points."""
    v, d = O.tabulate(cell, degree, 0, pts)
    d = d.reshape(len(pts), -1, M.TDIM[cell])
    return np.abs(v).sum(1).max(), np.abs(d).sum(1).max()


def scaled_tolerances(cell, p, pts, base, jac):
    """The reference's tolerance of each factor at degree p (module docstring)."""
    ref = 3 if cell == 7 else 4
    a_val, a_der = amplification(cell, p, pts)
    r_val, r_der = amplification(cell, ref, pts)
    n, n_ref = len(O.lagrange_points(cell, p)), len(O.lagrange_points(cell, ref))
    t_x = base * max(1.0, a_val / r_val)
    t_j = base * max(1.0, a_der / r_der) * max(1.0, n / n_ref)
    jm = jac.reshape(-1, 3, 3)  # [..][d][g]: columns of J
    kappa = (np.linalg.norm(jm, axis=(1, 2)) * np.linalg.norm(np.linalg.inv(jm), axis=(1, 2))).max()
    return {"points": t_x, "jacobians": t_j, "detj": t_j, "pinv": t_j * max(1.0, kappa / 3)}


@pytest.mark.parametrize("p", [5, 6, 7, 8])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_high_degree_tetrahedra(cuda, p, dtype):
    import torch

    mesh = volume_mesh(TET, p, n=2)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    g = api.geometry_from_host(mesh, dtype=tdt)
    ps = api.PointSet.gauss(TET, p + 1)
    names = ["points", "jacobians", "detj", "pinv"]
    want = oracle_factors(g, mesh, 0, ps, names)
    tol = scaled_tolerances(TET, p, ps.points, 1e-12 if dtype == "f64" else 1e-5,
                            want["jacobians"])
    for label, opt in (("default", 0), ("runtime-k", api.OPT_FORCE_GENERIC),
                       ("cooperative", api.OPT_COOPERATIVE)):
        err = rel_errs(gpu_factors(g, 0, ps, names, options=opt), want)
        print(f"P{p} tetrahedra {dtype} {label}: {err} (tolerances {tol})")
        for name, e in err.items():
            assert e <= tol[name], (label, name, e, tol[name])


def test_cooperative_runs_are_bitwise_identical(cuda):
    """The tree has a fixed shape: two runs give the same bytes."""
    import torch

    mesh = volume_mesh(TET, 7, n=2)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    ps = api.PointSet.gauss(TET, 8)
    a = gpu_factors(g, 0, ps, ["jacobians", "detj"])
    b = gpu_factors(g, 0, ps, ["jacobians", "detj"])
    for n in a:
        assert np.array_equal(a[n], b[n]), n
