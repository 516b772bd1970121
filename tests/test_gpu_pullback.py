"""tg_pull_back_block / pull_back_block and Mesh.pull_back (pull_back.rs:26-180, 566) on the GPU.

Pinned by round trips (the physical points of Gauss points come back to those points), by the
FP64 oracle's restatement of the same iteration (or_pull_back, with the oracle's own bases) for
points inside, slightly outside and far outside a cell, by the closest point of a surface for
points off it, and by affine cells, where one step is exact. Tolerances: FP64 1e-12 on xi
(oracle and round trips); FP32 2e-5 (round trips)."""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

IV, TRI, QUAD, TET, HEX, PRISM, PYR = (M.INTERVAL, M.TRIANGLE, M.QUADRILATERAL, M.TETRAHEDRON,
                                       M.HEXAHEDRON, M.PRISM, M.PYRAMID)


def volume_mesh(cell, p, n=2, seed=31):
    m = M.cube_box(cell, n, p)
    x = m.nodes
    bump = np.prod(np.sin(np.pi * x), axis=1)[:, None] * np.array([1.0, 0.6, 0.3])
    m = M.HostMesh(3, x + 0.12 / n * bump, m.blocks)
    return M.jitter(m, 0.03 / (n * p), seed, interior_only=True)


def planar_mesh(cell, p, gdim, n=3, seed=32):
    m = M.square_box(cell, n, p, gdim=gdim)
    if gdim == 3:
        x = m.nodes
        x[:, 2] = 0.2 * np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1])
    return M.jitter(m, 0.04 / (n * p), seed, interior_only=True)


def curve_mesh(p, gdim, n=6):
    """n intervals of degree p along [0, 1] (1D) or the arc (t, 0.1 sin(pi t)) (2D), each with
    its own nodes in A.5 order (vertices, then interior nodes)."""
    ref = M.reference_nodes(IV, p)[:, 0]
    t = ((np.arange(n)[:, None] + ref[None, :]) / n).reshape(-1)
    t = t + 0.2 / (n * p) * np.sin(3.0 * t) * (t > 0) * (t < 1)
    nodes = t[:, None] if gdim == 1 else np.stack([t, 0.1 * np.sin(np.pi * t)], axis=1)
    cells = np.arange(n * (p + 1), dtype=np.uint32).reshape(n, p + 1)
    return M.HostMesh(gdim, nodes, [M.HostBlock(IV, p, cells)])


def mesh_of(cell, p, gdim):
    if M.TDIM[cell] == 3:
        return volume_mesh(cell, p)
    if cell == IV:
        return curve_mesh(p, gdim)
    return planar_mesh(cell, p, gdim)


def to_field(x_cell_major, tile=16):
    """A QuadField (device, FP64) of the cell-major points [c][q][gdim]."""
    import torch

    c, nq, gd = x_cell_major.shape
    padded = (c + tile - 1) // tile * tile
    full = np.concatenate([x_cell_major, np.repeat(x_cell_major[-1:], padded - c, axis=0)])
    data = full.reshape(padded // tile, tile, nq, gd).transpose(0, 2, 3, 1).reshape(-1)
    return api.QuadField(torch.tensor(np.ascontiguousarray(data), device="cuda"), c, nq, gd, tile)


def torch_dtype(dtype):
    import torch

    return torch.float64 if dtype == "f64" else torch.float32


CASES = ([(IV, p, g) for p in (1, 2, 3, 4) for g in (1, 2)]
         + [(TET, p, 3) for p in (1, 2, 3, 4)] + [(HEX, p, 3) for p in (1, 2, 3)]
         + [(PRISM, p, 3) for p in (1, 2, 3, 4)] + [(PYR, 1, 3)]
         + [(TRI, p, g) for p in (1, 2, 3, 4) for g in (2, 3)]
         + [(QUAD, p, g) for p in (1, 2, 3, 4) for g in (2, 3)])


@pytest.mark.parametrize("cell,p,gdim", CASES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_round_trip_of_gauss_points(cuda, cell, p, gdim, dtype):
    """X(xi_q) of every cell, pulled back, gives xi_q again, converged (step <= tolerance)."""
    import torch

    mesh = mesh_of(cell, p, gdim)
    g = api.geometry_from_host(mesh, dtype=torch_dtype(dtype))
    ps = api.PointSet.gauss(cell, 2 * p)
    x = api.physical_points_block(g, 0, ps)
    xi, step = api.pull_back_block(g, 0, x)
    torch.cuda.synchronize()
    got = xi.to_host_cell_major()
    want = np.broadcast_to(ps.points[None], got.shape)
    tol, conv = (1e-12, 1e-13) if dtype == "f64" else (2e-5, 1e-5)
    assert np.abs(got - want).max() <= tol
    assert step.to_host_cell_major().max() <= conv


def oracle_pull_back(g, mesh, b, x_cell_major, max_it=16, tol=1e-13):
    blk = mesh.blocks[b]
    nodes = g.nodes.detach().cpu().numpy().astype(np.float64) + np.asarray(g.origin)[None, :]
    return O.pull_back(blk.cell, blk.degree, blk.family, g.gdim, nodes, blk.cells.reshape(-1),
                       x_cell_major, max_it, tol)


@pytest.mark.parametrize("cell,p", [(TET, 2), (HEX, 2), (PRISM, 2), (PYR, 1), (TET, 3), (TET, 4),
                                    (PRISM, 4)])
def test_matches_oracle_inside_and_outside(cuda, cell, p):
    """FP64, against the oracle's iteration: points inside the cells, points up to 0.05 outside
    in reference coordinates (their preimages lie in the box [-0.1, 1.1], so they converge to
    reference points outside the cell), and points far away (clamped iterates that do not
    converge: step above the tolerance in both)."""
    import torch

    mesh = volume_mesh(cell, p)
    g = api.geometry_from_host(mesh, dtype=torch.float64)
    rng = np.random.default_rng(7)
    inside = api.PointSet.gauss(cell, 3).points
    near = inside.copy()
    near[:, 0] = -0.05  # outside the cell, inside the box
    pts = np.concatenate([inside, near])
    ps = api.PointSet.custom(cell, pts, np.ones(len(pts)))
    x = api.physical_points_block(g, 0, ps)
    torch.cuda.synchronize()
    xh = x.to_host_cell_major()
    far = xh[:, :2] + rng.uniform(3.0, 5.0, size=xh[:, :2].shape)
    xh = np.concatenate([xh, far], axis=1)
    xi, step = api.pull_back_block(g, 0, to_field(xh))
    torch.cuda.synchronize()
    got_xi, got_st = xi.to_host_cell_major(), step.to_host_cell_major()[:, :, 0]
    want_xi, want_st = oracle_pull_back(g, mesh, 0, xh)
    n_conv = len(pts)
    assert np.abs(got_xi[:, :n_conv] - want_xi[:, :n_conv]).max() <= 1e-12
    assert np.abs(got_xi[:, :n_conv] - pts[None]).max() <= 1e-12
    assert (got_st[:, :n_conv] <= 1e-13).all() and (want_st[:, :n_conv] <= 1e-13).all()
    # Far points: not converged in either; their finite iterates stay in the box.
    assert (got_st[:, n_conv:] > 1e-13).all() and (want_st[:, n_conv:] > 1e-13).all()
    fin = got_xi[np.isfinite(got_xi)]
    assert fin.min() >= float(np.float32(-0.1)) and fin.max() <= float(np.float32(1.1))


@pytest.mark.parametrize("cell,p", [(TRI, 2), (QUAD, 2), (TRI, 3), (QUAD, 4)])
def test_surface_points_pull_back_to_closest_point(cuda, cell, p):
    """Points moved off a curved surface by 0.01 along its normal pull back (least squares) to
    the reference point they came from (the closest point of the surface), as in the oracle."""
    import torch

    mesh = planar_mesh(cell, p, 3)
    g = api.geometry_from_host(mesh, dtype=torch.float64)
    ps = api.PointSet.gauss(cell, 2 * p)
    f = api.geometric_factors_block(g, 0, ps, ["points", "normals"])
    torch.cuda.synchronize()
    xh = f["points"].to_host_cell_major() + 0.01 * f["normals"].to_host_cell_major()
    xi, step = api.pull_back_block(g, 0, to_field(xh))
    torch.cuda.synchronize()
    got = xi.to_host_cell_major()
    assert np.abs(got - ps.points[None]).max() <= 1e-10
    want_xi, _ = oracle_pull_back(g, mesh, 0, xh)
    assert np.abs(got - want_xi).max() <= 1e-12
    assert step.to_host_cell_major().max() <= 1e-13


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_affine_cells_take_one_step(cuda, dtype):
    """On straight-sided tetrahedra X is affine, so the first step lands on the answer and the
    second step is (to rounding) zero: two iterations suffice."""
    import torch

    mesh = M.jitter(M.cube_box(TET, 3, 1), 0.02, 9, interior_only=True)
    g = api.geometry_from_host(mesh, dtype=torch_dtype(dtype))
    ps = api.PointSet.gauss(TET, 4)
    x = api.physical_points_block(g, 0, ps)
    xi, step = api.pull_back_block(g, 0, x, max_iterations=2)
    torch.cuda.synchronize()
    tol = 1e-13 if dtype == "f64" else 1e-5
    assert step.to_host_cell_major().max() <= tol
    err = np.abs(xi.to_host_cell_major() - ps.points[None]).max()
    assert err <= (1e-13 if dtype == "f64" else 2e-6)


def test_tiles_and_mesh_level_and_errors(cuda):
    """The result does not depend on the fields' tile (bitwise); Mesh.pull_back covers every
    block of the mixed mesh; a field of another size is LengthMismatch, an element without a
    kernel (Q4 hexahedra, P5 tetrahedra) UnsupportedDegree, and an empty point set launches
    nothing."""
    import torch

    mesh = volume_mesh(TET, 2)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    ps = api.PointSet.gauss(TET, 4)
    ref = None
    for tile in (16, 4, 1):
        x = api.physical_points_block(g, 0, ps, tile=tile)
        xi, step = api.pull_back_block(g, 0, x)
        torch.cuda.synchronize()
        got = (xi.to_host_cell_major(), step.to_host_cell_major())
        if ref is None:
            ref = got
        else:
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])

    h = M.config4_mesh(seed=4, scale=0.002)
    gm = api.geometry_from_host(h, dtype=torch.float64)
    rule = api.QuadRule.for_grid(gm, 3)
    xs = api.Mesh(gm).physical_points(rule)
    pb = api.Mesh(gm).pull_back(xs)
    torch.cuda.synchronize()
    for b, blk in enumerate(h.blocks):
        want = rule.get(blk.cell).points
        got = pb.reference[b].to_host_cell_major()
        assert np.abs(got - want[None]).max() <= 1e-11, b
        assert pb.last_step[b].to_host_cell_major().max() <= 1e-13, b

    other = volume_mesh(TET, 1, n=3)
    g2 = api.geometry_from_host(other, dtype=torch.float32)
    x2 = api.physical_points_block(g2, 0, ps)
    with pytest.raises(api.LengthMismatch):
        api.pull_back_block(g, 0, x2)
    g4 = api.geometry_from_host(volume_mesh(HEX, 4, n=1), dtype=torch.float32)
    x4 = api.physical_points_block(g4, 0, api.PointSet.gauss(HEX, 4))
    with pytest.raises(api.UnsupportedDegree):
        api.pull_back_block(g4, 0, x4)
    g5 = api.geometry_from_host(volume_mesh(TET, 5, n=1), dtype=torch.float32)
    x5 = api.physical_points_block(g5, 0, ps)
    with pytest.raises(api.UnsupportedDegree):
        api.pull_back_block(g5, 0, x5)
    empty = api.PointSet.custom(TET, np.zeros((0, 3)), np.zeros(0))
    x0 = api.physical_points_block(g, 0, empty)
    xi0, _ = api.pull_back_block(g, 0, x0)
    assert xi0.num_points == 0
