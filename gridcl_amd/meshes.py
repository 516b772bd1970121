"""Synthetic meshes for the geometry path, with the conventions of the reference's test meshes.

``cube_box`` follows tetraloma-testing/src/meshes/boxes.rs:155-216 and lattice.rs:1-222: nodes on
an integer lattice of resolution 2·n·p, hexahedral subcells in (k, j, i) order with i fastest,
the Kuhn split into six tetrahedra per hexahedron along the main diagonal (the same in every
subcell, so face diagonals conform), two prisms per subcell split along (0,0)-(1,1), six
pyramids per subcell around its centre, positive orientation by swapping axis vertices, the
Lagrange nodes of each cell as the affine image of the equispaced reference nodes in exact
integer arithmetic, and nodes numbered in order of first use.

Randomness comes from one documented generator: splitmix64 (Steele, Lea, Flood 2014) with state
seed + (i + 1)·0x9E3779B97F4A7C15 for the i-th draw, mapped to [0, 1) as (z >> 11)·2^-53.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib

# tetraloma_dims::CellType order.
POINT, INTERVAL, TRIANGLE, QUADRILATERAL, TETRAHEDRON, HEXAHEDRON, PRISM, PYRAMID = range(8)
TDIM = {INTERVAL: 1, TRIANGLE: 2, QUADRILATERAL: 2, TETRAHEDRON: 3, HEXAHEDRON: 3, PRISM: 3,
        PYRAMID: 3}
# Vertices whose differences to vertex 0 are the axes of the affine map (lattice.rs:20-29).
AXIS_VERTICES = {INTERVAL: [1], TRIANGLE: [1, 2], QUADRILATERAL: [1, 2],
                 TETRAHEDRON: [1, 2, 3], PRISM: [1, 2, 3], HEXAHEDRON: [1, 2, 4],
                 PYRAMID: [1, 2, 4]}
HEX_FACES = [[0, 1, 2, 3], [0, 1, 4, 5], [0, 2, 4, 6], [1, 3, 5, 7], [2, 3, 6, 7],
             [4, 5, 6, 7]]

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1) from splitmix64 draws 0..n-1 of `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class HostBlock:
    cell: int
    degree: int
    cells: np.ndarray  # uint32 [num_cells][k]
    family: int = 0
    flags: np.ndarray | None = None  # uint32 [num_cells]

    @property
    def num_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def nodes_per_cell(self) -> int:
        return int(self.cells.shape[1])


@dataclass
class HostMesh:
    gdim: int
    nodes: np.ndarray  # float64 [num_nodes][gdim], absolute coordinates
    blocks: list[HostBlock] = field(default_factory=list)

    @property
    def num_nodes(self) -> int:
        return int(self.nodes.shape[0])


def reference_nodes(cell: int, degree: int, family: int = 0) -> np.ndarray:
    """The library's A.5 Lagrange nodes [k][tdim] (host code, no GPU needed)."""
    lib = _lib.load()
    size = ctypes.c_size_t()
    rc = lib.tg_lagrange_points(cell, degree, family, None, 0, ctypes.byref(size))
    if rc:
        raise ValueError(_lib.last_error())
    out = (ctypes.c_double * size.value)()
    rc = lib.tg_lagrange_points(cell, degree, family, out, size.value, ctypes.byref(size))
    if rc:
        raise ValueError(_lib.last_error())
    return np.frombuffer(out, dtype=np.float64).reshape(-1, TDIM[cell]).copy()


def _axis_det(axes: np.ndarray) -> np.ndarray:
    """Determinant of the axes [..., td, 3] restricted to their first td coordinates."""
    td = axes.shape[-2]
    m = axes[..., :td, :td].astype(np.float64)
    return np.linalg.det(m) if td > 1 else m[..., 0, 0]


def _orient(cell: int, verts: np.ndarray) -> np.ndarray:
    """Swaps axis vertices of negatively oriented cells (lattice.rs:56-78)."""
    ax = np.array(AXIS_VERTICES[cell])
    axes = verts[:, ax, :] - verts[:, :1, :]
    neg = _axis_det(axes) < 0
    if np.any(neg):
        verts = verts.copy()
        if cell in (TRIANGLE, TETRAHEDRON, PYRAMID):
            verts[neg, 1], verts[neg, 2] = verts[neg, 2].copy(), verts[neg, 1].copy()
        elif cell == PRISM:
            verts[neg, 1], verts[neg, 2] = verts[neg, 2].copy(), verts[neg, 1].copy()
            verts[neg, 4], verts[neg, 5] = verts[neg, 5].copy(), verts[neg, 4].copy()
        else:
            raise ValueError("negatively oriented tensor cell")
    return verts


def _assemble(cell: int, degree: int, verts: np.ndarray, resolution: int, gdim: int) -> HostMesh:
    """Nodes of every cell (affine images of the reference nodes on the lattice), deduplicated
    in order of first use (lattice.rs:190-222)."""
    td = TDIM[cell]
    ref = np.rint(reference_nodes(cell, degree) * degree).astype(np.int64)  # [k][td]
    ax = np.array(AXIS_VERTICES[cell])
    axes = verts[:, ax, :] - verts[:, :1, :]  # [C][td][3]
    num = np.einsum("kt,ctg->ckg", ref, axes)  # [C][k][3]
    if np.any(num % degree):
        raise AssertionError("node off the lattice")
    keys = verts[:, :1, :] + num // degree  # [C][k][3]
    r1 = resolution + 1
    code = (keys[..., 0] * r1 + keys[..., 1]) * r1 + keys[..., 2]
    flat = code.reshape(-1)
    uniq, first, inverse = np.unique(flat, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    new_id = np.empty(len(uniq), dtype=np.int64)
    new_id[order] = np.arange(len(uniq))
    dofmap = new_id[inverse].reshape(code.shape).astype(np.uint32)
    ucode = uniq[order]
    xyz = np.stack([ucode // (r1 * r1), (ucode // r1) % r1, ucode % r1], axis=1)
    nodes = xyz[:, :gdim].astype(np.float64) / resolution
    del td
    return HostMesh(gdim=gdim, nodes=nodes, blocks=[HostBlock(cell, degree, dofmap)])


def cube_box(cell: int, n: int, p: int) -> HostMesh:
    """The unit cube in n^3 hexahedral subcells of degree p, split into `cell`s
    (boxes.rs:155-216)."""
    if n < 1 or p < 1:
        raise ValueError("a box needs n >= 1 and p >= 1")
    res = 2 * n * p
    h = 2 * p
    kk, jj, ii = np.indices((n, n, n)).reshape(3, -1)
    o = np.stack([ii, jj, kk], axis=1).astype(np.int64) * h  # [S][3], i fastest
    bits = np.array([[(v >> b) & 1 for b in range(3)] for v in range(8)], dtype=np.int64)
    corners = o[:, None, :] + bits[None, :, :] * h  # [S][8][3]
    centre = o + p
    if cell == HEXAHEDRON:
        verts = corners
    elif cell == TETRAHEDRON:
        # Kuhn: the path 0 -> e_a -> e_a + e_b -> 7 (boxes.rs:164-174).
        tets = []
        for a, b in [(0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1)]:
            ea = 1 << a
            tets.append(corners[:, [0, ea, ea | (1 << b), 7], :])
        verts = np.stack(tets, axis=1).reshape(-1, 4, 3)
    elif cell == PRISM:
        pr = [corners[:, [a, b, d, a + 4, b + 4, d + 4], :] for a, b, d in [(0, 1, 3), (0, 3, 2)]]
        verts = np.stack(pr, axis=1).reshape(-1, 6, 3)
    elif cell == PYRAMID:
        py = [np.concatenate([corners[:, f, :], centre[:, None, :]], axis=1) for f in HEX_FACES]
        verts = np.stack(py, axis=1).reshape(-1, 5, 3)
    else:
        raise ValueError("cube_box makes 3D cells")
    verts = _orient(cell, verts)
    return _assemble(cell, p, verts, res, 3)


def square_box(cell: int, n: int, p: int, gdim: int = 2) -> HostMesh:
    """The unit square in n^2 subcells of degree p: quadrilaterals, or two triangles per subcell
    split along (0,0)-(1,1); embedded in gdim = 2 or 3 (z = 0)."""
    res = 2 * n * p
    h = 2 * p
    jj, ii = np.indices((n, n)).reshape(2, -1)
    o = np.stack([ii, jj, np.zeros_like(ii)], axis=1).astype(np.int64) * h
    bits = np.array([[v & 1, (v >> 1) & 1, 0] for v in range(4)], dtype=np.int64)
    corners = o[:, None, :] + bits[None, :, :] * h
    if cell == QUADRILATERAL:
        verts = corners
    elif cell == TRIANGLE:
        verts = np.stack([corners[:, [0, 1, 3], :], corners[:, [0, 3, 2], :]], axis=1)
        verts = verts.reshape(-1, 3, 3)
        verts = _orient(TRIANGLE, verts)
    else:
        raise ValueError("square_box makes 2D cells")
    return _assemble(cell, p, verts, res, gdim)


def jitter(mesh: HostMesh, amplitude: float, seed: int, interior_only: bool) -> HostMesh:
    """Moves node i by amplitude·(2u - 1) per coordinate, u the draws 3i..3i+2 of `seed`;
    with interior_only, nodes on the boundary of the unit box keep their place (draws are still
    consumed, so every node's draws are fixed by its index)."""
    n, g = mesh.nodes.shape
    u = splitmix64_uniform(seed, n * g).reshape(n, g)
    delta = amplitude * (2.0 * u - 1.0)
    if interior_only:
        eps = 1e-12
        inside = np.all((mesh.nodes > eps) & (mesh.nodes < 1.0 - eps), axis=1)
        delta[~inside] = 0.0
    return HostMesh(gdim=mesh.gdim, nodes=mesh.nodes + delta, blocks=mesh.blocks)


def p1_to_p2_tets(mesh: HostMesh) -> HostMesh:
    """Straight-sided P2 tetrahedra from P1 ones: vertex nodes keep their indices, each edge
    gets one node at its chord midpoint (A.5 edge order (0,1),(0,2),(0,3),(1,2),(1,3),(2,3)),
    edge nodes numbered after the vertices in order of first use."""
    b = mesh.blocks[0]
    assert b.cell == TETRAHEDRON and b.degree == 1
    v = b.cells.astype(np.int64)
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    a = np.stack([v[:, i] for i, _ in edges], axis=1)
    c = np.stack([v[:, j] for _, j in edges], axis=1)
    lo, hi = np.minimum(a, c), np.maximum(a, c)
    code = (lo * mesh.num_nodes + hi).reshape(-1)
    uniq, first, inverse = np.unique(code, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")
    new_id = np.empty(len(uniq), dtype=np.int64)
    new_id[order] = np.arange(len(uniq))
    edge_node = new_id[inverse].reshape(-1, 6) + mesh.num_nodes
    u = uniq[order]
    mids = 0.5 * (mesh.nodes[u // mesh.num_nodes] + mesh.nodes[u % mesh.num_nodes])
    nodes = np.concatenate([mesh.nodes, mids], axis=0)
    cells = np.concatenate([v, edge_node], axis=1).astype(np.uint32)
    return HostMesh(gdim=3, nodes=nodes, blocks=[HostBlock(TETRAHEDRON, 2, cells)])


# ---- The benchmark configurations --------------------------------------------------------------

def config1_mesh(seed: int = 1, n: int = 32) -> HostMesh:
    """P1 tetrahedra: n^3 hexahedra x 6 Kuhn tets, interior vertices jittered by +-0.1h."""
    m = cube_box(TETRAHEDRON, n, 1)
    return jitter(m, 0.1 / n, seed, interior_only=True)


def config2_mesh(seed: int = 2, n: int = 64) -> HostMesh:
    """Curved P2 tetrahedra: n^3 x 6, every node displaced by
    0.05h·(sin2πx sin2πy, sin2πy sin2πz, sin2πz sin2πx) plus uniform jitter +-0.02h."""
    m = cube_box(TETRAHEDRON, n, 2)
    h = 1.0 / n
    x = m.nodes
    s = np.sin(2.0 * np.pi * x)
    disp = 0.05 * h * np.stack([s[:, 0] * s[:, 1], s[:, 1] * s[:, 2], s[:, 2] * s[:, 0]], axis=1)
    m = HostMesh(gdim=3, nodes=x + disp, blocks=m.blocks)
    return jitter(m, 0.02 * h, seed, interior_only=False)


def config3_mesh(seed: int = 3, n: int = 96) -> HostMesh:
    """Q3 hexahedra: n^3 cells, nodes at xi + 0.08 sin(πxi_x) sin(πxi_y) sin(πxi_z)·(1,1,1)
    plus uniform jitter +-0.01h."""
    m = cube_box(HEXAHEDRON, n, 3)
    x = m.nodes
    bump = 0.08 * np.prod(np.sin(np.pi * x), axis=1)
    m = HostMesh(gdim=3, nodes=x + bump[:, None], blocks=m.blocks)
    return jitter(m, 0.01 / n, seed, interior_only=False)


def config5_mesh(seed: int = 1, n: int = 32) -> HostMesh:
    """The P1 mesh of config 1 as straight-sided P2 tetrahedra, then every interior node
    jittered by a further +-0.02h (seed + 100) so that the cells are curved."""
    m = p1_to_p2_tets(config1_mesh(seed, n))
    return jitter(m, 0.02 / n, seed + 100, interior_only=True)


def p3tet_mesh(seed: int = 6, n: int = 48) -> HostMesh:
    """Curved P3 tetrahedra: n^3 x 6, config 2's displacement field (amplitude 0.05h) and
    uniform jitter +-0.02h of every node, at degree 3 (20 nodes per cell)."""
    m = cube_box(TETRAHEDRON, n, 3)
    h = 1.0 / n
    x = m.nodes
    s = np.sin(2.0 * np.pi * x)
    disp = 0.05 * h * np.stack([s[:, 0] * s[:, 1], s[:, 1] * s[:, 2], s[:, 2] * s[:, 0]], axis=1)
    m = HostMesh(gdim=3, nodes=x + disp, blocks=m.blocks)
    return jitter(m, 0.02 * h, seed, interior_only=False)


HYBRID_NOMINAL = [
    # (cell, degree, nominal cells)
    (TETRAHEDRON, 2, 600_000),
    (HEXAHEDRON, 2, 250_000),
    (PRISM, 2, 180_000),
    (PYRAMID, 1, 40_000),
    (TRIANGLE, 3, 300_000),
]


def config4_mesh(seed: int = 4, scale: float = 1.0) -> HostMesh:
    """The hybrid mesh: one block per row of HYBRID_NOMINAL, each of size nominal·scale·(1 + 0.6u
    - 0.3) with u the block's draw of `seed`, cut from its own box; volume boxes jittered at
    interior nodes by +-0.1 h/p (config 1's +-0.1h at the node spacing h/p of degree p); the P3
    triangles are the unit square lifted to the
    surface z = 0.1 sin(πx) sin(πy) in 3D, every other cell flagged (bit 0)."""
    u = splitmix64_uniform(seed, len(HYBRID_NOMINAL))
    nodes, blocks = [], []
    offset = 0
    per_subcell = {TETRAHEDRON: 6, HEXAHEDRON: 1, PRISM: 2, PYRAMID: 6, TRIANGLE: 2}
    for (cell, p, nominal), ub in zip(HYBRID_NOMINAL, u):
        target = max(1, int(round(nominal * scale * (0.7 + 0.6 * ub))))
        if cell == TRIANGLE:
            n = int(np.ceil(np.sqrt(target / 2.0)))
            m = square_box(TRIANGLE, n, p, gdim=3)
            x = m.nodes
            x[:, 2] = 0.1 * np.sin(np.pi * x[:, 0]) * np.sin(np.pi * x[:, 1])
        else:
            n = int(np.ceil((target / per_subcell[cell]) ** (1.0 / 3.0)))
            m = cube_box(cell, n, p)
            m = jitter(m, 0.1 / (n * max(1, p)), seed + 17 * (cell + 1), interior_only=True)
        b = m.blocks[0]
        cells = b.cells[:target].astype(np.int64) + offset
        flags = None
        if cell == TRIANGLE:
            flags = (np.arange(target) % 2).astype(np.uint32)
        blocks.append(HostBlock(cell, p, cells.astype(np.uint32), flags=flags))
        nodes.append(m.nodes)
        offset += m.num_nodes
    return HostMesh(gdim=3, nodes=np.concatenate(nodes, axis=0), blocks=blocks)
