"""tg_equispaced_nodes / equispaced_nodes (equispaced.rs:1-12, 158) on the GPU.

Pinned by: the same geometry (GLL and equispaced nodes of one degree interpolate the same
polynomial map, so the converted nodes read as an equispaced block give the GLL block's X at any
point, FP64 1e-13); the first-occurrence rule, repeated on the host from the GPU's own X at the
lattice, listed lexicographically like the conversion does (bitwise); the FP64 oracle's GLL basis at the equispaced lattice (1e-13); equispaced
blocks and unused nodes copied unchanged (bitwise); the same result on every run."""
import numpy as np
import pytest

from gridcl_amd import api
from gridcl_amd import meshes as M
from oracle import oracle as O

pytestmark = pytest.mark.gpu

IV, QUAD, HEX = M.INTERVAL, M.QUADRILATERAL, M.HEXAHEDRON
GLL, EQ = 1, 0


def gll_box(cell, n, p, gdim, seed=41):
    """The unit box in n^tdim cells of degree p with GLL nodes shared between neighbours,
    mapped by a smooth curving field (and lifted into gdim > tdim)."""
    tdim = M.TDIM[cell]
    ref = M.reference_nodes(cell, p, GLL)  # [k][tdim]
    g1 = np.unique(np.round(ref[:, 0], 12))
    local = np.stack([np.abs(ref[:, d][:, None] - g1[None, :]).argmin(axis=1)
                      for d in range(tdim)], axis=1)  # [k][tdim] lattice offsets
    L = n * p + 1
    cells = []
    for idx in np.ndindex(*([n] * tdim)):
        cell_idx = idx[::-1]  # first direction fastest
        gi = [(cell_idx[d] * p + local[:, d]) for d in range(tdim)]
        flat = sum(gi[d] * L ** d for d in range(tdim))
        cells.append(flat)
    cells = np.array(cells, dtype=np.uint32)
    t1 = np.concatenate([(i + g1[:-1]) / n for i in range(n)] + [[1.0]])
    grids = np.meshgrid(*([t1] * tdim), indexing="ij")
    t = np.stack([gr.transpose(*range(tdim)[::-1]).reshape(-1) for gr in grids[::-1]], axis=1)
    t = t[:, ::-1] if tdim > 1 else t
    bump = np.prod(np.sin(np.pi * t), axis=1)
    x = t + 0.08 * bump[:, None]
    if gdim > tdim:
        x = np.concatenate([x, 0.1 * bump[:, None]], axis=1)
    return M.HostMesh(gdim, np.ascontiguousarray(x), [M.HostBlock(cell, p, cells, family=GLL)])


CASES = [(IV, 3, 1), (IV, 4, 2), (QUAD, 2, 2), (QUAD, 3, 2), (QUAD, 4, 3), (HEX, 2, 3),
         (HEX, 3, 3)]


@pytest.mark.parametrize("cell,p,gdim", CASES)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_equispaced_conversion(cuda, cell, p, gdim, dtype):
    import torch

    tdt = torch.float64 if dtype == "f64" else torch.float32
    mesh = gll_box(cell, 3, p, gdim)
    g = api.geometry_from_host(mesh, dtype=tdt)
    out = api.equispaced_nodes(g)
    again = api.equispaced_nodes(g)
    torch.cuda.synchronize()
    assert torch.equal(out, again)
    # The first-occurrence rule from the GPU's own X at the lattice, relative to the origin. The
    # conversion lists a tensor cell's lattice lexicographically (first coordinate fastest, so
    # that hexahedra take the sum-factorised kernel): X at that point set, read in A.5 order.
    a5 = M.reference_nodes(cell, p, EQ)
    lex = np.rint(a5 * p).astype(int) @ (p + 1) ** np.arange(a5.shape[1])
    pts = np.empty_like(a5)
    pts[lex] = a5
    lat = api.PointSet.custom(cell, pts, np.ones(len(pts)))
    f = api.geometric_factors_block(g, 0, lat, ["points"], options=api.OPT_POINTS_RELATIVE)
    torch.cuda.synchronize()
    xl = f["points"].to_host_cell_major()[:, lex]  # [c][k][gdim], A.5 order
    want = g.nodes.cpu().numpy().copy()
    seen = set()
    for c, row in enumerate(mesh.blocks[0].cells):
        for kk, node in enumerate(row):
            if node not in seen:
                seen.add(node)
                want[node] = xl[c, kk]
    assert np.array_equal(out.cpu().numpy(), want)
    # FP64: the same map read through equispaced nodes, and the oracle's GLL basis.
    if dtype == "f64":
        org = np.asarray(g.origin)[None, :]
        eq = M.HostMesh(gdim, out.cpu().numpy() + org,
                        [M.HostBlock(cell, p, mesh.blocks[0].cells, family=EQ)])
        g2 = api.geometry_from_host(eq, dtype=tdt)
        rng = np.random.default_rng(5)
        pts = rng.uniform(0.0, 1.0, size=(7, M.TDIM[cell]))
        ps = api.PointSet.custom(cell, pts, np.ones(len(pts)))
        a = api.physical_points_block(g, 0, ps).to_host_cell_major()
        b = api.physical_points_block(g2, 0, ps).to_host_cell_major()
        assert np.abs(a - b).max() <= 1e-13
        o = O.eval_block(cell, p, GLL, gdim, mesh.nodes, mesh.blocks[0].cells.reshape(-1),
                         M.reference_nodes(cell, p, EQ), factors=("points",))["points"]
        assert np.abs(xl + org[None] - o).max() <= 1e-13


def test_equispaced_blocks_and_unused_nodes_unchanged(cuda):
    """A GLL block and an equispaced block on separate nodes, plus a node no cell uses: only the
    GLL block's nodes change."""
    import torch

    gm = gll_box(QUAD, 2, 3, 2)
    eq = M.square_box(QUAD, 2, 2, gdim=2)
    n0 = gm.num_nodes
    nodes = np.concatenate([gm.nodes, eq.nodes + 2.0, [[5.0, 5.0]]])
    blocks = [gm.blocks[0], M.HostBlock(QUAD, 2, eq.blocks[0].cells + n0, family=EQ)]
    mesh = M.HostMesh(2, nodes, blocks)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    out = api.equispaced_nodes(g).cpu().numpy()
    before = g.nodes.cpu().numpy()
    assert np.array_equal(out[n0:], before[n0:])
    assert not np.array_equal(out[:n0], before[:n0])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_equispaced_second_call_after_nodes_move(cuda, dtype):
    """The first call's table of first occurrences (kept with the geometry) serves a later call
    on another stream after the nodes moved: doubling every node (exact in floating point)
    doubles every converted node bit for bit, as the conversion is linear in the nodes."""
    import torch

    tdt = torch.float64 if dtype == "f64" else torch.float32
    g = api.geometry_from_host(gll_box(HEX, 3, 3, 3), dtype=tdt)
    first = api.equispaced_nodes(g).clone()
    g.nodes.mul_(2.0)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        second = api.equispaced_nodes(g)
    side.synchronize()
    assert torch.equal(second, 2.0 * first)


def lattice_x(g, b, cell, p):
    """X (relative to the origin) of block b at its cells' equispaced lattice, [c][k][gdim] in A.5
    order, through the lexicographic point set the conversion uses."""
    import torch

    a5 = M.reference_nodes(cell, p, EQ)
    lex = np.rint(a5 * p).astype(int) @ (p + 1) ** np.arange(a5.shape[1])
    pts = np.empty_like(a5)
    pts[lex] = a5
    lat = api.PointSet.custom(cell, pts, np.ones(len(pts)))
    f = api.geometric_factors_block(g, b, lat, ["points"], options=api.OPT_POINTS_RELATIVE)
    torch.cuda.synchronize()
    return f["points"].to_host_cell_major()[:, lex]


@pytest.mark.parametrize("cell,p,gdim", [(QUAD, 3, 2), (HEX, 3, 3)])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_equispaced_two_gll_blocks_sharing_nodes(cuda, cell, p, gdim, dtype):
    """One GLL box cut into two blocks that share the nodes of the cut, the second half of the
    cells first: each shared node takes its value from its first occurrence in (block, cell,
    local node) order, i.e. from block 0, through the cached table's per-block start offsets.
    Bitwise against that rule repeated on the host from each block's X at the lattice, then again
    on another stream after the nodes moved (doubled: the conversion is linear, so bit for bit
    twice the first result)."""
    import torch

    tdt = torch.float64 if dtype == "f64" else torch.float32
    box = gll_box(cell, 4 if cell == QUAD else 3, p, gdim)
    cells = box.blocks[0].cells
    half = len(cells) // 2
    blocks = [M.HostBlock(cell, p, np.ascontiguousarray(cells[half:]), family=GLL),
              M.HostBlock(cell, p, np.ascontiguousarray(cells[:half]), family=GLL)]
    mesh = M.HostMesh(gdim, box.nodes, blocks)
    g = api.geometry_from_host(mesh, dtype=tdt)
    out = api.equispaced_nodes(g)
    torch.cuda.synchronize()
    want = g.nodes.cpu().numpy().copy()
    seen = set()
    shared = set(np.unique(blocks[0].cells)) & set(np.unique(blocks[1].cells))
    assert shared, "the two blocks share the nodes of the cut"
    for b, blk in enumerate(mesh.blocks):
        xl = lattice_x(g, b, cell, p)
        for c, row in enumerate(blk.cells):
            for kk, node in enumerate(row):
                if node not in seen:
                    seen.add(node)
                    want[node] = xl[c, kk]
    assert np.array_equal(out.cpu().numpy(), want)
    first = out.clone()
    g.nodes.mul_(2.0)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        second = api.equispaced_nodes(g)
    side.synchronize()
    assert torch.equal(second, 2.0 * first)
