"""Pins the FP64 oracle's factors and adjoint without any reference output (CPU only): affine
closed forms, exact volumes, K J = I (tetraloma-geometry/tests/factors.rs:51-185,
tetraloma/tests/unit_volume.rs), the tiled CPU path against the cell-major one, and the
functional's gradient against central differences and a hand-derived single-cell gradient."""
import numpy as np
import pytest

from oracle import oracle as O
from gridcl_amd import meshes as M
from gridcl_amd.api import PointSet

TET, HEX, PRISM = 4, 5, 6


def test_affine_tets_closed_form():
    mesh = M.config1_mesh(n=2)
    blk = mesh.blocks[0]
    ps = PointSet.tet_14()
    f = O.eval_block(TET, 1, 0, 3, mesh.nodes, blk.cells.reshape(-1), ps.points,
                     factors=("jacobians", "detj", "pinv", "points"))
    x = mesh.nodes[blk.cells.astype(np.int64)]
    cols = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=1)
    assert np.abs(f["jacobians"] - cols.reshape(-1, 1, 9)).max() <= 1e-15
    det = np.linalg.det(cols.transpose(0, 2, 1))
    assert np.abs(f["detj"][:, :, 0] - det[:, None]).max() <= 1e-16
    # X is the affine image of the reference point.
    want = x[:, :1, :] + np.einsum("qd,cdg->cqg", ps.points, cols)
    assert np.abs(f["points"] - want).max() <= 1e-15
    # K J = I (factors.rs:51-147): K column-major tdim x gdim at g*3+t, J at d*3+g.
    K = f["pinv"].reshape(-1, 3, 3)   # [.., g, t]
    J = f["jacobians"].reshape(-1, 3, 3)  # [.., d, g]
    KJ = np.einsum("ngt,ndg->ntd", K, J)
    assert np.abs(KJ - np.eye(3)).max() <= 1e-12


@pytest.mark.parametrize("cell,p,deg", [(TET, 2, 3), (TET, 3, 6), (HEX, 2, 5), (PRISM, 2, 5),
                                        (HEX, 3, 8)])
def test_unit_volume(cell, p, deg):
    mesh = M.jitter(M.cube_box(cell, 3, p), 0.1 / (3 * p), 7, interior_only=True)
    blk = mesh.blocks[0]
    ps = PointSet.gauss(cell, deg)
    d = O.eval_block(cell, p, 0, 3, mesh.nodes, blk.cells.reshape(-1), ps.points,
                     factors=("detj",))["detj"][:, :, 0]
    assert abs((d @ ps.weights).sum() - 1.0) <= 1e-13


def test_p3_tetrahedron_mesh_is_valid():
    """The curved P3 mesh of the P3 benchmark leg: 6 n^3 cells of 20 nodes, det J > 0 at every
    Gauss point of degree 6, and a volume within what the jitter of the boundary nodes (+-0.02h
    on each of the six faces) can move from the unit cube's."""
    mesh = M.p3tet_mesh(n=3)
    blk = mesh.blocks[0]
    assert blk.cells.shape == (6 * 27, 20) and blk.degree == 3
    ps = PointSet.gauss(TET, 6)
    d = O.eval_block(TET, 3, 0, 3, mesh.nodes, blk.cells.reshape(-1), ps.points,
                     factors=("detj",))["detj"][:, :, 0]
    assert d.min() > 0
    assert abs((d @ ps.weights).sum() - 1.0) <= 6 * 0.02 / 3


def test_tiled_cpu_path_matches_cell_major():
    mesh = M.config2_mesh(n=3)
    blk = mesh.blocks[0]
    ps = PointSet.tet_14()
    origin = np.array([0.5, 0.5, 0.5])
    v, d = O.tabulate(TET, 2, 0, ps.points)
    x, j, dj = O.cpu_factors(3, 3, 10, mesh.nodes - origin, origin, blk.cells, v, d, tile=4)
    want = O.eval_block(TET, 2, 0, 3, mesh.nodes, blk.cells.reshape(-1), ps.points)
    nc, nq = blk.num_cells, len(ps)
    pad = (nc + 3) // 4 * 4

    def untile(a, length):
        a = a.reshape(pad // 4, nq, length, 4).transpose(0, 3, 1, 2)
        return a.reshape(pad, nq, length)[:nc]

    # The tiled path works on coordinates relative to the origin, the cell-major one on absolute
    # coordinates: the two differ by rounding only.
    assert O.rel_err(untile(x, 3), want["points"]) <= 1e-15
    assert O.rel_err(untile(j, 9), want["jacobians"]) <= 1e-13
    assert O.rel_err(untile(dj, 1), want["detj"]) <= 1e-13


def test_functional_gradient_central_differences():
    mesh = M.config5_mesh(n=2)
    blk = mesh.blocks[0]
    ps = PointSet.tet_14()
    F, g = O.functional(TET, 2, 0, mesh.nodes, blk.cells.reshape(-1), ps.points, ps.weights,
                        beta=0.1)
    h = 1e-6
    for i in range(16):
        d = M.splitmix64_uniform(50 + i, mesh.nodes.size).reshape(mesh.nodes.shape) * 2 - 1
        fp, _ = O.functional(TET, 2, 0, mesh.nodes + h * d, blk.cells.reshape(-1), ps.points,
                             ps.weights, beta=0.1, grad=False)
        fm, _ = O.functional(TET, 2, 0, mesh.nodes - h * d, blk.cells.reshape(-1), ps.points,
                             ps.weights, beta=0.1, grad=False)
        fd = (fp - fm) / (2 * h)
        an = float((g * d).sum())
        assert abs(fd - an) <= 1e-7 * abs(an), i


def test_functional_single_affine_tet():
    """One affine tetrahedron: F = vol + beta vol |J|^2 / det^(2/3) exactly, and with beta = 0
    dF/dx_k = cof(J) dphi_k / 6 (dphi_0 = -(1,1,1), dphi_i = e_i)."""
    x = np.array([[0.1, 0.0, 0.0], [1.2, 0.1, 0.0], [0.2, 0.9, 0.1], [0.0, 0.2, 1.1]])
    cells = np.array([0, 1, 2, 3], dtype=np.uint32)
    ps = PointSet.tet_keast4()
    J = np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)  # rows g, columns d
    det = np.linalg.det(J)
    F, g = O.functional(TET, 1, 0, x, cells, ps.points, ps.weights, beta=0.1)
    want = det / 6 + 0.1 * (J ** 2).sum() / det ** (2 / 3) / 6
    assert abs(F - want) <= 1e-15 * abs(want)
    F0, g0 = O.functional(TET, 1, 0, x, cells, ps.points, ps.weights, beta=0.0)
    cof = det * np.linalg.inv(J).T  # cof[g, d]
    dphi = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    want_g = (dphi @ cof.T) / 6  # [k][g] = sum_d cof[g, d] dphi[k, d] / 6
    assert np.abs(g0 - want_g).max() <= 1e-15


@pytest.mark.parametrize("cell,p", [(TET, 1), (TET, 2), (TET, 3), (TET, 4), (HEX, 2), (PRISM, 2),
                                    (PRISM, 4), (7, 1)])
def test_pull_back_round_trip(cell, p):
    """or_pull_back (pull_back.rs:26-180) inverts the oracle's own X: the physical points of
    reference points come back within 1e-13, converged (last step <= 1e-13)."""
    mesh = M.jitter(M.cube_box(cell, 2, p), 0.05 / (2 * p), 3, interior_only=True)
    blk = mesh.blocks[0]
    pts = np.array([[0.1, 0.2, 0.15], [0.2, 0.2, 0.2], [0.05, 0.3, 0.4], [0.3, 0.1, 0.05]])
    x = O.eval_block(cell, p, 0, 3, mesh.nodes, blk.cells.reshape(-1), pts,
                     factors=("points",))["points"]
    xi, step = O.pull_back(cell, p, 0, 3, mesh.nodes, blk.cells.reshape(-1), x)
    assert np.abs(xi - pts[None]).max() <= 1e-13
    assert step.max() <= 1e-13


def test_pull_back_affine_one_step_and_far_points():
    """Affine tetrahedra: after one step the next is zero, so two iterations converge; a point
    far outside never converges and its iterates are clamped to [-0.1f, 1.1f]."""
    mesh = M.config1_mesh(n=2)
    blk = mesh.blocks[0]
    pts = PointSet.tet_14().points
    x = O.eval_block(TET, 1, 0, 3, mesh.nodes, blk.cells.reshape(-1), pts,
                     factors=("points",))["points"]
    xi, step = O.pull_back(TET, 1, 0, 3, mesh.nodes, blk.cells.reshape(-1), x, max_it=2)
    assert np.abs(xi - pts[None]).max() <= 1e-14 and step.max() <= 1e-13
    far = x + 10.0
    xi, step = O.pull_back(TET, 1, 0, 3, mesh.nodes, blk.cells.reshape(-1), far)
    assert (step > 1e-13).all()
    assert xi.min() >= float(np.float32(-0.1)) and xi.max() <= float(np.float32(1.1))


def test_pull_back_surface_closest_point():
    """A point at distance 0.01 along the normal of a curved surface pulls back (least squares)
    to the reference point of its foot: the residual x - X(xi) is normal to both tangents."""
    m = M.square_box(2, 2, 2, gdim=3)
    m.nodes[:, 2] = 0.2 * np.sin(np.pi * m.nodes[:, 0]) * np.sin(np.pi * m.nodes[:, 1])
    blk = m.blocks[0]
    pts = np.array([[0.2, 0.3], [0.6, 0.1], [0.1, 0.1]])
    f = O.eval_block(2, 2, 0, 3, m.nodes, blk.cells.reshape(-1), pts,
                     factors=("points", "normals", "jacobians"))
    x = f["points"] + 0.01 * f["normals"]
    xi, step = O.pull_back(2, 2, 0, 3, m.nodes, blk.cells.reshape(-1), x)
    assert np.abs(xi - pts[None]).max() <= 1e-12 and step.max() <= 1e-13
    J = f["jacobians"].reshape(*f["jacobians"].shape[:2], 2, 3)
    r = x - f["points"]
    assert np.abs(np.einsum("cqdg,cqg->cqd", J, r)).max() <= 1e-15


@pytest.mark.parametrize("p", [1, 2, 3])
def test_cpu_port_matches_oracle(p):
    """The CPU-baseline port (the reference's CPU schedule: E-vector, then X, J and det J passes
    over tiles of 4 cells, multiply-adds fused) gives the oracle's values to FP64 rounding."""
    mesh = M.jitter(M.cube_box(TET, 3, p), 0.05 / p, 7, interior_only=True)
    blk = mesh.blocks[0]
    ps = PointSet.tet_keast4() if p == 1 else PointSet.tet_14()
    origin = np.array([0.5, 0.5, 0.5])
    values, derivs = O.tabulate(TET, p, 0, ps.points)
    port = O.CpuPort(blk.nodes_per_cell, mesh.nodes - origin, origin, blk.cells, values, derivs,
                     threads=2)
    x, j, dj = port()
    want = O.eval_block(TET, p, 0, 3, mesh.nodes, blk.cells.reshape(-1), ps.points)
    c, nq = blk.num_cells, len(ps)
    pad = (c + 3) // 4 * 4

    def untile(a, ln):
        return a.reshape(pad // 4, nq, ln, 4).transpose(0, 3, 1, 2).reshape(pad, nq, ln)[:c]

    assert O.rel_err(untile(x, 3), want["points"]) < 1e-14
    assert O.rel_err(untile(j, 9), want["jacobians"]) < 1e-13
    assert O.rel_err(untile(dj, 1), want["detj"]) < 1e-13
