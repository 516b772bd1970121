"""Python mirror of the reference's geometry API over the C ABI (include/tetraloma_geom.h).

Names, argument meanings and errors follow tetraloma-geometry's public API (src/lib.rs:78-99):
``Geometry`` (geometry.rs:49), ``GeometryBlock`` / ``CoordinateElement`` (element.rs:30-70),
``PointSet`` (point_set.rs:57), ``QuadRule`` (rule.rs:49), ``FacetQuadRule`` (facet_rule.rs:204),
``QuadField`` (field.rs:162), the block functions ``physical_points_block``, ``jacobians_block``
(points_jac.rs:42, 90), ``detj_block``, ``pseudo_inverse_block`` (detj_pinv.rs:42, 90),
``facet_normals_block``, ``normals_block`` (facets.rs:46, 107) and the mesh-level loops of
``Mesh`` (tetraloma/src/mesh.rs:249-257, 434-964). ``geometric_factors_block`` is the fused
evaluation of several factors in one kernel pass, which the reference does not have.

Device memory is held in torch tensors on the current HIP device; every evaluation is enqueued
on torch's current stream. Nothing here computes geometry on the host.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from enum import IntEnum
from typing import Iterable, Sequence

import numpy as np
import torch

from . import _lib


class CellType(IntEnum):
    POINT = 0
    INTERVAL = 1
    TRIANGLE = 2
    QUADRILATERAL = 3
    TETRAHEDRON = 4
    HEXAHEDRON = 5
    PRISM = 6
    PYRAMID = 7


class NodeFamily(IntEnum):
    EQUISPACED = 0
    GLL = 1


TDIM = {CellType.POINT: 0, CellType.INTERVAL: 1, CellType.TRIANGLE: 2,
        CellType.QUADRILATERAL: 2, CellType.TETRAHEDRON: 3, CellType.HEXAHEDRON: 3,
        CellType.PRISM: 3, CellType.PYRAMID: 3}

# Factor bits of the C ABI and their names.
FACTORS = {"points": 1, "jacobians": 2, "detj": 4, "pinv": 8, "facet_normals": 16, "normals": 32}

OPT_NO_CELL_SHIFT = 1
OPT_NO_FP64_FIXUP = 2
OPT_FORCE_GENERIC = 4
OPT_MAP_CELL = 8
OPT_MAP_CELL_STAGED = 16
OPT_MAP_TILE = 32
OPT_NO_SUMFACT = 64
OPT_PLAN_PER_BLOCK = 128
OPT_PLAN_SERIAL = 256
OPT_SUMFACT_ALWAYS = 512
OPT_POINTS_RELATIVE = 1024
OPT_PLAN_PERSISTENT = 2048
OPT_NO_MFMA = 4096
OPT_COOPERATIVE = 8192


# ---- Errors: one class per tetraloma_device::GridError variant of the path -------------------

class GridError(Exception):
    code = -1


class LengthMismatch(GridError):
    code = 1


class TooLarge(GridError):
    code = 2


class AllocationTooLarge(GridError):
    code = 3


class UnsupportedDegree(GridError):
    code = 4


class UnsupportedNodeFamily(GridError):
    code = 5


class CellNotInRule(GridError):
    code = 6


class UnsupportedPrecision(GridError):
    code = 7


class InvalidArgument(GridError):
    code = 8


class ServerError(GridError):
    code = 9


# The Gmsh reader's variants (tetraloma-device/src/error.rs:129-190). `detail` is the number the
# variant carries: the repeated id, the element tag, the byte offset or the element type.
class DimensionMismatch(GridError):
    code = 10


class DuplicateOriginId(GridError):
    code = 11


class DuplicateCellBlock(GridError):
    code = 12


class UnmatchedTagElement(GridError):
    code = 13


class Io(GridError):
    code = 14


class Parse(GridError):
    code = 15


class UnsupportedElementType(GridError):
    code = 16


class IndexOutOfRange(GridError):
    code = 17


_ERRORS = {c.code: c for c in (LengthMismatch, TooLarge, AllocationTooLarge, UnsupportedDegree,
                               UnsupportedNodeFamily, CellNotInRule, UnsupportedPrecision,
                               InvalidArgument, ServerError, DimensionMismatch, DuplicateOriginId,
                               DuplicateCellBlock, UnmatchedTagElement, Io, Parse,
                               UnsupportedElementType, IndexOutOfRange)}


def _check(rc: int) -> None:
    if rc != 0:
        raise _ERRORS.get(rc, GridError)(_lib.last_error())


def _dptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _dbl(a: np.ndarray | None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a


# ---- Elements, blocks, geometry -----------------------------------------------------------------

@dataclass(frozen=True)
class CoordinateElement:
    """The coordinate element of a block (element.rs:30-62)."""

    cell: CellType
    degree: int
    family: NodeFamily = NodeFamily.EQUISPACED

    @property
    def tdim(self) -> int:
        return TDIM[CellType(self.cell)]

    @property
    def nodes_per_cell(self) -> int:
        return int(_lib.load().tg_num_lagrange_nodes(int(self.cell), int(self.degree)))


@dataclass
class GeometryBlock:
    """One cell block: its element, its dofmap (u32 [cells][k], held as a device int32 tensor of
    the same bits) and optional cell flags (bit 0 flips normals)."""

    element: CoordinateElement
    dofmap: torch.Tensor
    flags: torch.Tensor | None = None
    stride: int | None = None  # defaults to the element's node count


def _u32_device(a: np.ndarray | torch.Tensor, device) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        if a.dtype not in (torch.int32,):
            raise TypeError("dofmaps and flags are int32 tensors holding u32 bits")
        return a.contiguous()
    arr = np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)
    return torch.from_numpy(arr.copy()).to(device)


class Geometry:
    """Node coordinates relative to an origin, plus cell blocks (geometry.rs:49-150).

    ``nodes`` is a device tensor [num_nodes][gdim] of float32 or float64 (the geometry's
    precision); ``origin`` the gdim f64 values that X adds back."""

    def __init__(self, blocks: Sequence[GeometryBlock], nodes: torch.Tensor,
                 origin: Sequence[float]):
        lib = _lib.load()
        if nodes.dim() != 2:
            raise LengthMismatch("nodes must be [num_nodes][gdim]")
        if nodes.dtype == torch.float32:
            self.dtype = 0
        elif nodes.dtype == torch.float64:
            self.dtype = 1
        else:
            raise UnsupportedPrecision(f"no geometry of {nodes.dtype}")
        self.nodes = nodes.contiguous()
        self.gdim = int(nodes.shape[1])
        self.origin = [float(o) for o in origin]
        self.blocks = list(blocks)
        descs = (_lib.BlockDesc * max(1, len(self.blocks)))()
        for i, b in enumerate(self.blocks):
            k = b.element.nodes_per_cell
            descs[i].cell = int(b.element.cell)
            descs[i].degree = int(b.element.degree)
            descs[i].family = int(b.element.family)
            descs[i].dofmap = _dptr(b.dofmap) if b.dofmap.numel() else None
            descs[i].dofmap_len = int(b.dofmap.numel())
            descs[i].dofmap_stride = int(b.stride if b.stride is not None else k)
            descs[i].flags = _dptr(b.flags) if b.flags is not None else None
        org = (ctypes.c_double * 3)(*(self.origin + [0.0] * (3 - len(self.origin))))
        handle = ctypes.c_void_p()
        _check(lib.tg_geometry_create(self.dtype, self.gdim,
                                      _dptr(self.nodes) if self.nodes.numel() else None,
                                      int(self.nodes.shape[0]), org, descs, len(self.blocks),
                                      ctypes.byref(handle)))
        self._h = handle

    @property
    def handle(self):
        return self._h

    def num_blocks(self) -> int:
        return len(self.blocks)

    def block_cells(self, b: int) -> int:
        return int(_lib.load().tg_geometry_block_cells(self._h, b))

    @property
    def torch_dtype(self):
        return torch.float32 if self.dtype == 0 else torch.float64

    def set_fixup_threshold(self, tau: float) -> None:
        _check(_lib.load().tg_geometry_set_fixup_threshold(self._h, float(tau)))

    def set_nodes_f64(self, nodes_f64: torch.Tensor | None) -> None:
        """FP32 geometries: the node coordinates also in FP64 (device [num_nodes][gdim], relative
        to the same origin), which the evaluations gather so that each cell's offsets are taken
        in FP64 and rounded once (tg_geometry_set_nodes_f64); None detaches."""
        if nodes_f64 is not None:
            if nodes_f64.dtype != torch.float64 or tuple(nodes_f64.shape) != tuple(self.nodes.shape):
                raise LengthMismatch("FP64 nodes must be float64 of the geometry's node shape")
            nodes_f64 = nodes_f64.contiguous()
        self.nodes_f64 = nodes_f64
        ptr = _dptr(nodes_f64) if nodes_f64 is not None and nodes_f64.numel() else None
        _check(_lib.load().tg_geometry_set_nodes_f64(self._h, ptr))

    def evector(self, b: int) -> torch.Tensor:
        """The E-vector of block b, [k][gdim][cells padded to 16] (evector.rs:197-236)."""
        lib = _lib.load()
        n = int(lib.tg_evector_elems(self._h, b))
        out = torch.empty(max(n, 1), dtype=self.torch_dtype, device=self.nodes.device)
        _check(lib.tg_evector_build(self._h, b, _dptr(out), _stream()))
        return out[:n]

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().tg_geometry_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- Point sets and rules -----------------------------------------------------------------------

class PointSet:
    """Points [nq][tdim] and weights of one cell type (point_set.rs:57-110), optionally with the
    outward reference normal of each point's facet (a facet point set, facet_rule.rs:39)."""

    def __init__(self, handle, cell: CellType, facet: bool = False):
        self._h = handle
        self.cell = CellType(cell)
        self.is_facet = facet
        lib = _lib.load()
        nq = int(lib.tg_pointset_len(handle))
        td = TDIM[self.cell]
        pts = np.zeros(max(1, nq * td))
        wts = np.zeros(max(1, nq))
        nrm = np.zeros(max(1, nq * td))
        count = ctypes.c_uint64()
        dp = ctypes.POINTER(ctypes.c_double)
        _check(lib.tg_pointset_get(handle, pts.ctypes.data_as(dp), wts.ctypes.data_as(dp),
                                   nrm.ctypes.data_as(dp), ctypes.byref(count)))
        self.points = pts[: nq * td].reshape(nq, td)
        self.weights = wts[:nq]
        self.facet_normals = nrm[: nq * td].reshape(nq, td) if count.value else None

    @classmethod
    def custom(cls, cell: CellType, points, weights, facet_normals=None) -> "PointSet":
        lib = _lib.load()
        p = _dbl(np.asarray(points, dtype=np.float64).reshape(-1))
        w = _dbl(np.asarray(weights, dtype=np.float64).reshape(-1))
        n = _dbl(None if facet_normals is None else np.asarray(facet_normals).reshape(-1))
        h = ctypes.c_void_p()
        _check(lib.tg_pointset_create(int(cell), p[0], w[0], len(w[1]),
                                      n[0] if n else None, ctypes.byref(h)))
        return cls(h, cell, facet=facet_normals is not None)

    @classmethod
    def gauss(cls, cell: CellType, degree: int) -> "PointSet":
        h = ctypes.c_void_p()
        _check(_lib.load().tg_pointset_gauss(int(cell), int(degree), ctypes.byref(h)))
        return cls(h, cell)

    @classmethod
    def facet_gauss(cls, cell: CellType, degree: int) -> "PointSet":
        h = ctypes.c_void_p()
        _check(_lib.load().tg_pointset_facet_gauss(int(cell), int(degree), ctypes.byref(h)))
        return cls(h, cell, facet=True)

    @classmethod
    def tet_keast4(cls) -> "PointSet":
        h = ctypes.c_void_p()
        _check(_lib.load().tg_pointset_tet_rule(0, ctypes.byref(h)))
        return cls(h, CellType.TETRAHEDRON)

    @classmethod
    def tet_14(cls) -> "PointSet":
        h = ctypes.c_void_p()
        _check(_lib.load().tg_pointset_tet_rule(1, ctypes.byref(h)))
        return cls(h, CellType.TETRAHEDRON)

    def __len__(self) -> int:
        return len(self.weights)

    def tabulate(self, e: CoordinateElement) -> np.ndarray:
        """phi [q][k] in f64 (point_set.rs:312)."""
        return self._tab(e)[0]

    def tabulate_derivatives(self, e: CoordinateElement) -> np.ndarray:
        """dphi [q][k][tdim] in f64 (point_set.rs:350)."""
        return self._tab(e)[1]

    def _tab(self, e: CoordinateElement):
        if CellType(e.cell) != self.cell:
            raise CellNotInRule(self.cell.name)
        k = e.nodes_per_cell
        nq, td = len(self), TDIM[self.cell]
        v = np.zeros(max(1, nq * k))
        d = np.zeros(max(1, nq * k * td))
        _check(_lib.load().tg_pointset_tabulate(
            self._h, int(e.degree), int(e.family),
            v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            d.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return v[: nq * k].reshape(nq, k), d[: nq * k * td].reshape(nq, k, td)

    def __del__(self):
        try:
            if self._h:
                _lib.load().tg_pointset_destroy(self._h)
                self._h = None
        except Exception:
            pass


class QuadRule:
    """One point set per cell type (rule.rs:49-172)."""

    def __init__(self, sets: Iterable[PointSet]):
        self.sets = list(sets)

    @classmethod
    def gauss(cls, cells: Iterable[CellType], degree: int) -> "QuadRule":
        seen = []
        for c in cells:
            if CellType(c) not in seen:
                seen.append(CellType(c))
        return cls(PointSet.gauss(c, degree) for c in seen)

    @classmethod
    def for_grid(cls, g: Geometry, degree: int) -> "QuadRule":
        return cls.gauss([b.element.cell for b in g.blocks], degree)

    def get(self, cell: CellType) -> PointSet:
        for s in self.sets:
            if s.cell == CellType(cell):
                return s
        raise CellNotInRule(CellType(cell).name)


class FacetQuadRule(QuadRule):
    """Facet point sets per cell type (facet_rule.rs:204-310)."""

    @classmethod
    def gauss(cls, cells: Iterable[CellType], degree: int) -> "FacetQuadRule":
        seen = []
        for c in cells:
            if CellType(c) not in seen:
                seen.append(CellType(c))
        return cls(PointSet.facet_gauss(c, degree) for c in seen)


# ---- Fields ------------------------------------------------------------------------------------

class QuadField:
    """Values at the points of one block in tiles of `tile` cells (field.rs:126-168): component
    k at point q of cell c at ((c // tile)·nq·LEN + q·LEN + k)·tile + c % tile."""

    def __init__(self, data: torch.Tensor, num_cells: int, num_points: int, length: int,
                 tile: int):
        self.data = data
        self.num_cells = num_cells
        self.num_points = num_points
        self.len = length
        self.tile = tile

    def to_host_cell_major(self) -> np.ndarray:
        """[cells][points][LEN] on the host, tiling undone (field.rs:380)."""
        t = self.tile
        a = self.data.detach().cpu().numpy()
        padded = (self.num_cells + t - 1) // t * t
        a = a[: padded * self.num_points * self.len]
        a = a.reshape(padded // t, self.num_points, self.len, t).transpose(0, 3, 1, 2)
        return a.reshape(padded, self.num_points, self.len)[: self.num_cells]


def factor_len(name: str, tdim: int, gdim: int) -> int:
    return {"points": gdim, "jacobians": tdim * gdim, "detj": 1, "pinv": tdim * gdim,
            "facet_normals": gdim, "normals": gdim}[name]


def geometric_factors_block(g: Geometry, b: int, pts: PointSet, factors: Sequence[str],
                            tile: int = 16, options: int = 0,
                            out: dict | None = None) -> dict[str, QuadField]:
    """Every factor in `factors` of block b at `pts`, in one fused kernel pass. `out` may
    supply preallocated device tensors by factor name."""
    lib = _lib.load()
    if b < 0 or b >= len(g.blocks):
        raise InvalidArgument("block index out of range")
    blk = g.blocks[b]
    td = TDIM[CellType(blk.element.cell)]
    ncells = g.block_cells(b)
    nq = len(pts)
    mask = 0
    fields = {}
    outs = _lib.Outputs()
    for name in factors:
        if name not in FACTORS:
            raise InvalidArgument(f"unknown factor {name}")
        mask |= FACTORS[name]
        length = factor_len(name, td, g.gdim)
        n = int(lib.tg_field_elems(ncells, nq, length, tile))
        if tile not in (1, 2, 4, 8, 16):
            raise InvalidArgument("tile must be 1, 2, 4, 8 or 16")
        if out is not None and name in out:
            data = out[name]
            if data.numel() < n or data.dtype != g.torch_dtype:
                raise LengthMismatch(f"output {name}: need {n} elements of {g.torch_dtype}")
        else:
            data = torch.empty(max(n, 1), dtype=g.torch_dtype, device=g.nodes.device)
        setattr(outs, name, _dptr(data))
        fields[name] = QuadField(data, ncells, nq, length, tile)
    _check(lib.tg_eval_block(g.handle, b, pts._h, mask, tile, options, ctypes.byref(outs),
                             _stream()))
    return fields


def physical_points_block(g: Geometry, b: int, pts: PointSet, tile: int = 16) -> QuadField:
    return geometric_factors_block(g, b, pts, ["points"], tile)["points"]


def jacobians_block(g: Geometry, b: int, pts: PointSet, tile: int = 16) -> QuadField:
    return geometric_factors_block(g, b, pts, ["jacobians"], tile)["jacobians"]


def detj_block(g: Geometry, b: int, pts: PointSet, tile: int = 16) -> QuadField:
    return geometric_factors_block(g, b, pts, ["detj"], tile)["detj"]


def pseudo_inverse_block(g: Geometry, b: int, pts: PointSet, tile: int = 16) -> QuadField:
    return geometric_factors_block(g, b, pts, ["pinv"], tile)["pinv"]


def facet_normals_block(g: Geometry, b: int, pts: PointSet, tile: int = 16) -> QuadField:
    return geometric_factors_block(g, b, pts, ["facet_normals"], tile)["facet_normals"]


def normals_block(g: Geometry, b: int, pts: PointSet, tile: int = 16) -> QuadField:
    return geometric_factors_block(g, b, pts, ["normals"], tile)["normals"]


def measure_block(g: Geometry, b: int, pts: PointSet, options: int = 0) -> float:
    """The length, area or volume of block b, sum_c sum_q w_q det J (measure.rs:141): the
    reference's fixed-order sums in the geometry's type on the device, widened to FP64 and read
    back once (the call synchronises, as the reference's read does)."""
    val = torch.empty((), dtype=torch.float64, device="cuda")
    _check(_lib.load().tg_measure_block(g.handle, b, pts._h, options, val.data_ptr(), _stream()))
    return float(val.item())


def equispaced_nodes(g: Geometry) -> torch.Tensor:
    """The node array with every GLL block's nodes moved to the equispaced nodes of the same
    geometry (equispaced.rs:158): [num_nodes][gdim], relative to the origin like g.nodes."""
    out = torch.empty_like(g.nodes)
    _check(_lib.load().tg_equispaced_nodes(g.handle, _dptr(out), _stream()))
    return out


def pull_back_block(g: Geometry, b: int, x: QuadField, max_iterations: int = 16,
                    tolerance: float | None = None, options: int = 0
                    ) -> tuple[QuadField, QuadField]:
    """The reference points of the physical points x (a field of block b with gdim components,
    absolute coordinates) and each point's last Newton step, as fields of x's tiling
    (pull_back.rs:566). tolerance None: 1e-13 in FP64, 1e-5 in FP32 (pull_back.rs:49-51)."""
    lib = _lib.load()
    if b < 0 or b >= len(g.blocks):
        raise InvalidArgument("block index out of range")
    ncells = g.block_cells(b)
    if x.num_cells != ncells:  # pull_back.rs:578-583
        raise LengthMismatch(f"LengthMismatch: expected {ncells} cells, found {x.num_cells}")
    if x.len != g.gdim or x.data.dtype != g.torch_dtype:
        raise InvalidArgument("x must be a field of gdim components in the geometry's type")
    td = TDIM[CellType(g.blocks[b].element.cell)]
    nq, tile = x.num_points, x.tile
    n_xi = int(lib.tg_field_elems(ncells, nq, td, tile))
    n_st = int(lib.tg_field_elems(ncells, nq, 1, tile))
    xi = torch.empty(max(n_xi, 1), dtype=g.torch_dtype, device=g.nodes.device)
    st = torch.empty(max(n_st, 1), dtype=g.torch_dtype, device=g.nodes.device)
    tol = -1.0 if tolerance is None else float(tolerance)
    _check(lib.tg_pull_back_block(g.handle, b, _dptr(x.data), nq, tile, int(max_iterations), tol,
                                  options, _dptr(xi), _dptr(st), _stream()))
    return QuadField(xi, ncells, nq, td, tile), QuadField(st, ncells, nq, 1, tile)


class PullBack:
    """Mesh.pull_back's result (pull_back.rs:62-99): per block, the reference points and the
    size of each point's last step."""

    def __init__(self, reference: list[QuadField], last_step: list[QuadField]):
        self.reference = reference
        self.last_step = last_step


class MeshPlan:
    """Every block of a geometry evaluated by one plan: one launch per block (the default;
    replayed from a hipGraph its blocks run as parallel branches) or, with OPT_PLAN_PERSISTENT,
    one persistent launch over a work list of all blocks.
    ``fields[b][name]`` are the outputs of block b, filled by ``launch()``/``graph_launch()``."""

    def __init__(self, g: Geometry, rule: QuadRule, factors: Sequence[str], tile: int = 16,
                 options: int = 0):
        lib = _lib.load()
        self.g = g
        self.sets = [rule.get(b.element.cell) for b in g.blocks]
        nb = len(g.blocks)
        outs = (_lib.Outputs * max(1, nb))()
        handles = (ctypes.c_void_p * max(1, nb))()
        mask = 0
        for name in factors:
            mask |= FACTORS[name]
        self.fields: list[dict[str, QuadField]] = []
        for b, blk in enumerate(g.blocks):
            td = TDIM[CellType(blk.element.cell)]
            ps = self.sets[b]
            handles[b] = ps._h
            allowed = {"points", "jacobians", "detj", "pinv"}
            if td == g.gdim and ps.is_facet:
                allowed.add("facet_normals")
            if td + 1 == g.gdim:
                allowed.add("normals")
            fb = {}
            ncells = g.block_cells(b)
            for name in factors:
                if name not in allowed:
                    continue
                length = factor_len(name, td, g.gdim)
                n = int(lib.tg_field_elems(ncells, len(ps), length, tile))
                data = torch.empty(max(n, 1), dtype=g.torch_dtype, device=g.nodes.device)
                setattr(outs[b], name, _dptr(data))
                fb[name] = QuadField(data, ncells, len(ps), length, tile)
            self.fields.append(fb)
        h = ctypes.c_void_p()
        _check(lib.tg_plan_create(g.handle, handles, mask, tile, options, outs, ctypes.byref(h)))
        self._h = h
        items, grid, pers = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32()
        _check(lib.tg_plan_info(h, ctypes.byref(items), ctypes.byref(grid), ctypes.byref(pers)))
        self.num_items, self.grid, self.persistent = items.value, grid.value, bool(pers.value)

    def launch(self) -> None:
        _check(_lib.load().tg_plan_launch(self._h, _stream()))

    def graph_launch(self) -> None:
        _check(_lib.load().tg_plan_graph_launch(self._h, _stream()))

    def __del__(self):
        try:
            if self._h:
                _lib.load().tg_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass


class Mesh:
    """The block loop of tetraloma::Mesh's geometric methods (mesh.rs:249-257, 434-964)."""

    def __init__(self, geometry: Geometry):
        self.geometry = geometry

    def geometric_factors(self, rule: QuadRule, factors: Sequence[str],
                          tile: int = 16) -> list[dict[str, QuadField]]:
        """Every block's factors in one launch (a MeshPlan launched once)."""
        plan = MeshPlan(self.geometry, rule, factors, tile)
        plan.launch()
        return plan.fields

    def _per_block(self, rule: QuadRule, name: str, tile: int) -> list[QuadField]:
        g = self.geometry
        return [geometric_factors_block(g, b, rule.get(blk.element.cell), [name], tile)[name]
                for b, blk in enumerate(g.blocks)]

    def physical_points(self, rule: QuadRule, tile: int = 16):
        return self._per_block(rule, "points", tile)

    def jacobians(self, rule: QuadRule, tile: int = 16):
        return self._per_block(rule, "jacobians", tile)

    def detj(self, rule: QuadRule, tile: int = 16):
        return self._per_block(rule, "detj", tile)

    def pseudo_inverse_jacobians(self, rule: QuadRule, tile: int = 16):
        return self._per_block(rule, "pinv", tile)

    def normals(self, rule: QuadRule, tile: int = 16):
        return self._per_block(rule, "normals", tile)

    def facet_normals(self, rule: FacetQuadRule, tile: int = 16):
        return self._per_block(rule, "facet_normals", tile)

    def pull_back(self, x: list[QuadField], max_iterations: int = 16,
                  tolerance: float | None = None) -> PullBack:
        """pull_back_block of every block's points (one field per block)."""
        g = self.geometry
        if len(x) != len(g.blocks):
            raise LengthMismatch(f"LengthMismatch: expected {len(g.blocks)} fields, found {len(x)}")
        out = [pull_back_block(g, b, f, max_iterations, tolerance) for b, f in enumerate(x)]
        return PullBack([o[0] for o in out], [o[1] for o in out])

    def measure(self, degree: int) -> float:
        """The sum over blocks of measure_block at QuadRule.for_grid(degree) (mesh.rs:579-586;
        one rank, so no reduction over ranks)."""
        g = self.geometry
        rule = QuadRule.for_grid(g, degree)
        local = 0.0
        for b, blk in enumerate(g.blocks):
            local += measure_block(g, b, rule.get(blk.element.cell))
        return local


class MeshFunctional:
    """F(X) = sum_c sum_q w_q det J + beta sum_c sum_q w_q |J|_F^2 / det(J)^(2/3) over block b
    at `pts`, with its vector-Jacobian product (no reference counterpart).

    ``forward(nodes)`` returns F as a 0-d FP64 device tensor; ``vjp(seed, nodes)`` returns
    seed * dF/dX as an FP64 device tensor [num_nodes][3]; ``apply(nodes)`` is the same pair as a
    torch.autograd.Function, so ``F = fn.apply(x); F.backward()`` fills ``x.grad``."""

    def __init__(self, g: Geometry, b: int, pts: PointSet, beta: float = 0.1):
        self.g, self.b, self.pts, self.beta = g, b, pts, float(beta)
        h = ctypes.c_void_p()
        _check(_lib.load().tg_functional_create(g.handle, b, pts._h, ctypes.byref(h)))
        self._h = h

    def _nodes(self, nodes):
        if nodes is None:
            return None
        if nodes.dtype != self.g.torch_dtype or nodes.shape != self.g.nodes.shape:
            raise LengthMismatch("nodes must match the geometry's shape and type")
        return _dptr(nodes.contiguous())

    def forward(self, nodes: torch.Tensor | None = None) -> torch.Tensor:
        out = torch.empty((), dtype=torch.float64, device=self.g.nodes.device)
        _check(_lib.load().tg_functional_forward(self._h, self._nodes(nodes), self.beta,
                                                 _dptr(out), _stream()))
        return out

    def vjp(self, seed: torch.Tensor | None = None,
            nodes: torch.Tensor | None = None) -> torch.Tensor:
        grad = torch.empty(self.g.nodes.shape, dtype=torch.float64, device=self.g.nodes.device)
        s = None
        if seed is not None:
            seed = seed.to(dtype=torch.float64).reshape(()).contiguous()
            s = _dptr(seed)
        _check(_lib.load().tg_functional_vjp(self._h, self._nodes(nodes), self.beta, s,
                                             _dptr(grad), _stream()))
        return grad

    def apply(self, nodes: torch.Tensor) -> torch.Tensor:
        return _FunctionalFn.apply(nodes, self)

    def __del__(self):
        try:
            if self._h:
                _lib.load().tg_functional_destroy(self._h)
                self._h = None
        except Exception:
            pass


class _FunctionalFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, nodes, fn):
        ctx.save_for_backward(nodes)
        ctx.fn = fn
        return fn.forward(nodes.detach())

    @staticmethod
    def backward(ctx, grad_out):
        (nodes,) = ctx.saved_tensors
        g = ctx.fn.vjp(grad_out, nodes.detach())
        return g.to(nodes.dtype), None


def geometry_from_host(mesh, dtype=torch.float32, device="cuda", origin=None, nodes_f64=False):
    """Uploads a HostMesh: nodes relative to the centre of their bounding box (the reference's
    origin, geometry.rs:14-19), narrowed to `dtype`; dofmaps and flags as u32 bits. nodes_f64
    (FP32 only): also upload the coordinates in FP64 for the evaluations (set_nodes_f64)."""
    x = np.asarray(mesh.nodes, dtype=np.float64)
    if origin is None:
        origin = 0.5 * (x.min(axis=0) + x.max(axis=0)) if len(x) else np.zeros(mesh.gdim)
    rel = x - np.asarray(origin)[None, :]
    np_dtype = np.float32 if dtype == torch.float32 else np.float64
    nodes = torch.from_numpy(rel.astype(np_dtype)).to(device)
    blocks = []
    for b in mesh.blocks:
        e = CoordinateElement(CellType(b.cell), int(b.degree), NodeFamily(b.family))
        dm = _u32_device(np.asarray(b.cells).reshape(-1), device)
        fl = None if b.flags is None else _u32_device(np.asarray(b.flags), device)
        blocks.append(GeometryBlock(e, dm, fl))
    g = Geometry(blocks, nodes, [float(o) for o in origin])
    if nodes_f64:
        g.set_nodes_f64(torch.from_numpy(rel).to(device))
    return g


def stored_nodes_f64(g: Geometry) -> np.ndarray:
    """The coordinates the geometry stores, widened to f64 plus the origin: the node values a
    host reference evaluates (tetraloma-testing/src/reference.rs:47-69)."""
    x = g.nodes.detach().cpu().numpy().astype(np.float64)
    return x + np.asarray(g.origin)[None, :]
