"""Gmsh MSH 4.1 files: the mesh a file describes, read by the library's host reader.

Mirrors ``tetraloma_io::gmsh::{read_file, parse, to_gmsh_mesh}`` (tetraloma-io/src/gmsh/
parse.rs:571, 621; convert.rs:229) and its ``GmshMesh`` / ``GmshCellBlock`` / ``GmshTagGroup``
(convert.rs:36-137): ASCII and binary files in either byte order and with 4- or 8-byte size_t;
cell blocks of the highest-dimensional elements in cell-type order with their nodes in A.5 order
(u32 rows into ``nodes``), serendipity cells completed to P2, lower-dimensional elements as tag
groups, physical names. Parsing and conversion run in C++ (``tg_gmsh_*``); this module copies the
handle's arrays into numpy and hands the cell blocks to the device through ``geometry``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import api
from . import meshes


@dataclass
class GmshCellBlock:
    cell: int
    degree: int
    nodes: np.ndarray       # uint32 [num_cells][k], A.5 order
    tags: np.ndarray        # int64 element tag per cell
    elementary: np.ndarray  # int32 entity tag per cell
    physical: np.ndarray    # int64 [n][2]: (cell index, physical tag), sorted


@dataclass
class GmshTagGroup:
    dim: int
    cell: int
    corners: np.ndarray     # int64 [num_elements][vertices], node tags, internal vertex order
    tags: np.ndarray
    elementary: np.ndarray
    physical: np.ndarray


@dataclass
class GmshMesh:
    tdim: int
    gdim: int
    nodes: np.ndarray       # float64 [num_nodes][gdim]
    node_tags: np.ndarray   # int64
    blocks: list[GmshCellBlock] = field(default_factory=list)
    tag_groups: list[GmshTagGroup] = field(default_factory=list)
    names: list[tuple[int, int, str]] = field(default_factory=list)
    # What the file held as stored.
    binary: bool = False
    file_nodes: int = 0
    element_blocks: list[tuple[int, int, int, int]] = field(default_factory=list)  # dim, entity, type, count
    partitioned: bool = False
    num_partitions: int = 0
    num_periodic: int = 0

    @property
    def element_counts(self) -> dict[int, int]:
        """Elements per Gmsh type id, as stored in the file."""
        out: dict[int, int] = {}
        for _, _, ty, n in self.element_blocks:
            out[ty] = out.get(ty, 0) + n
        return out

    def host_mesh(self) -> meshes.HostMesh:
        """The cell blocks as equispaced Lagrange blocks over the nodes."""
        return meshes.HostMesh(self.gdim, self.nodes,
                               [meshes.HostBlock(b.cell, b.degree, b.nodes, family=0)
                                for b in self.blocks])

    def geometry(self, dtype=None, device="cuda"):
        """A device geometry of the cell blocks (api.geometry_from_host)."""
        import torch

        return api.geometry_from_host(self.host_mesh(), dtype=dtype or torch.float32,
                                      device=device)


def _raise(rc: int) -> None:
    lib = _lib.load()
    cls = api._ERRORS.get(rc, api.GridError)
    err = cls(_lib.last_error())
    err.detail = int(lib.tg_gmsh_last_error_detail())
    raise err


def _array(ptr, dtype, count: int) -> np.ndarray:
    if count == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    buf = (ctypes.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype).copy()


def _from_handle(h) -> GmshMesh:
    lib = _lib.load()
    info = _lib.GmshInfo()
    if lib.tg_gmsh_get_info(h, ctypes.byref(info)):
        _raise(api.InvalidArgument.code)
    n = int(info.num_nodes)
    mesh = GmshMesh(
        tdim=int(info.tdim), gdim=int(info.gdim),
        nodes=_array(lib.tg_gmsh_nodes(h), np.float64, n * info.gdim).reshape(n, info.gdim),
        node_tags=_array(lib.tg_gmsh_node_tags(h), np.int64, n),
        binary=bool(info.binary), file_nodes=int(info.file_nodes),
        partitioned=bool(info.partitioned), num_partitions=int(info.num_partitions),
        num_periodic=int(info.num_periodic))
    for b in range(info.num_blocks):
        d = _lib.GmshBlock()
        lib.tg_gmsh_get_block(h, b, ctypes.byref(d))
        c, k = int(d.num_cells), int(d.nodes_per_cell)
        mesh.blocks.append(GmshCellBlock(
            int(d.cell), int(d.degree), _array(d.nodes, np.uint32, c * k).reshape(c, k),
            _array(d.tags, np.int64, c), _array(d.elementary, np.int32, c),
            _array(d.physical, np.int64, 2 * d.num_physical).reshape(-1, 2)))
    for g in range(info.num_tag_groups):
        d = _lib.GmshTagGroup()
        lib.tg_gmsh_get_tag_group(h, g, ctypes.byref(d))
        e, v = int(d.num_elements), int(d.corners_per_element)
        mesh.tag_groups.append(GmshTagGroup(
            int(d.dim), int(d.cell), _array(d.corners, np.int64, e * v).reshape(e, v),
            _array(d.tags, np.int64, e), _array(d.elementary, np.int32, e),
            _array(d.physical, np.int64, 2 * d.num_physical).reshape(-1, 2)))
    for i in range(info.num_names):
        dim, tag, name = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_char_p()
        lib.tg_gmsh_get_name(h, i, ctypes.byref(dim), ctypes.byref(tag), ctypes.byref(name))
        mesh.names.append((dim.value, tag.value, name.value.decode(errors="replace")))
    for i in range(info.num_element_blocks):
        d = _lib.GmshElementBlock()
        lib.tg_gmsh_get_element_block(h, i, ctypes.byref(d))
        mesh.element_blocks.append((int(d.dim), int(d.entity_tag), int(d.element_type),
                                    int(d.count)))
    return mesh


def _finish(rc: int, h: ctypes.c_void_p) -> GmshMesh:
    if rc:
        _raise(rc)
    try:
        return _from_handle(h)
    finally:
        _lib.load().tg_gmsh_destroy(h)


def read(path: str | os.PathLike, gdim: int | None = None) -> GmshMesh:
    """The mesh of the file at `path` (read_file + to_gmsh_mesh); `gdim` fixes D_GEO."""
    h = ctypes.c_void_p()
    rc = _lib.load().tg_gmsh_read_file(os.fsencode(path), gdim or 0, ctypes.byref(h))
    return _finish(rc, h)


def parse(data: bytes, gdim: int | None = None) -> GmshMesh:
    """The mesh of the file bytes `data` (parse + to_gmsh_mesh)."""
    h = ctypes.c_void_p()
    rc = _lib.load().tg_gmsh_read_bytes(data, len(data), gdim or 0, ctypes.byref(h))
    return _finish(rc, h)


def gmsh_to_internal(cell: int, degree: int) -> np.ndarray:
    """map[k] = the A.5 index of Gmsh's k-th node (tetraloma-refcell/src/node_maps.rs:54)."""
    lib = _lib.load()
    n = ctypes.c_size_t()
    rc = lib.tg_gmsh_to_internal(cell, degree, None, 0, ctypes.byref(n))
    if rc:
        _raise(rc)
    out = (ctypes.c_int32 * n.value)()
    rc = lib.tg_gmsh_to_internal(cell, degree, out, n.value, ctypes.byref(n))
    if rc:
        _raise(rc)
    return np.frombuffer(out, dtype=np.int32).copy()


# ---- Encoding files ----------------------------------------------------------------------------


@dataclass
class Sections:
    """An MSH 4.1 file's sections as stored (tetraloma-io/src/gmsh/model.rs): what ``encode``
    writes. Entities are dicts with keys tag, min, max, physical, boundary (and, for partitioned
    entities, parent_dim, parent_tag, partitions); periodic links have dim, tag, master, affine,
    nodes."""

    names: list = field(default_factory=list)             # (dim, tag, name)
    entities: list | None = None                          # 4 lists, by dimension
    partitioned: dict | None = None                       # num_partitions, ghosts, entities
    periodic: list = field(default_factory=list)
    node_blocks: list = field(default_factory=list)       # (dim, entity, tags[n], coords[n][3])
    element_blocks: list = field(default_factory=list)    # (dim, entity, type, tags[n], nodes[n][k])


class _Out:
    def __init__(self, binary: bool, big_endian: bool, size_t: int):
        self.binary, self.size_t = binary, size_t
        self.end = ">" if big_endian else "<"
        self.parts: list[bytes] = []

    def text(self, s: str) -> None:
        self.parts.append(s.encode())

    def ints(self, v) -> None:
        v = np.asarray(v, dtype=np.int64).reshape(-1)
        if self.binary:
            self.parts.append(v.astype(self.end + "i4").tobytes())
        else:
            self.text(" ".join(str(int(x)) for x in v) + "\n" if len(v) else "")

    def sizes(self, v) -> None:
        v = np.asarray(v, dtype=np.uint64).reshape(-1)
        if self.binary:
            self.parts.append(v.astype(self.end + ("u8" if self.size_t == 8 else "u4")).tobytes())
        else:
            self.text(" ".join(str(int(x)) for x in v) + "\n" if len(v) else "")

    def floats(self, v) -> None:
        v = np.asarray(v, dtype=np.float64).reshape(-1)
        if self.binary:
            self.parts.append(v.astype(self.end + "f8").tobytes())
        else:
            self.text(" ".join(repr(float(x)) for x in v) + "\n" if len(v) else "")


def _entity(o: _Out, e: dict, dim: int, partitioned: bool) -> None:
    o.ints([e["tag"]])
    if partitioned:
        o.ints([e["parent_dim"], e["parent_tag"]])
        o.sizes([len(e["partitions"])])
        o.ints(e["partitions"])
    o.floats(e["min"] if dim == 0 else list(e["min"]) + list(e["max"]))
    o.sizes([len(e["physical"])])
    o.ints(e["physical"])
    if dim > 0:
        o.sizes([len(e["boundary"])])
        o.ints(e["boundary"])


def encode(sec: Sections, binary: bool = True, big_endian: bool = False, size_t: int = 8) -> bytes:
    """The bytes of an MSH 4.1 file holding `sec`: ASCII, or binary in either byte order with a
    4- or 8-byte size_t (the sections and value types of Gmsh's format, as parse.rs reads them).
    Used to build test and benchmark inputs; it writes the sections as given and does not apply
    the reference writer's numbering rules (convert.rs:361-376)."""
    o = _Out(binary, big_endian, size_t)
    o.text(f"$MeshFormat\n4.1 {1 if binary else 0} {size_t}\n")
    if binary:
        o.parts.append(np.array([1], dtype=o.end + "i4").tobytes())
        o.text("\n")
    o.text("$EndMeshFormat\n")
    if sec.names:
        o.text(f"$PhysicalNames\n{len(sec.names)}\n")
        for dim, tag, name in sec.names:
            o.text(f'{dim} {tag} "{name}"\n')
        o.text("$EndPhysicalNames\n")
    if sec.entities is not None:
        o.text("$Entities\n")
        o.sizes([len(sec.entities[d]) for d in range(4)])
        for d in range(4):
            for e in sec.entities[d]:
                _entity(o, e, d, False)
        o.text("\n$EndEntities\n" if binary else "$EndEntities\n")
    if sec.partitioned is not None:
        p = sec.partitioned
        o.text("$PartitionedEntities\n")
        o.sizes([p["num_partitions"]])
        o.sizes([len(p["ghosts"])])
        for g in p["ghosts"]:
            o.ints(g)
        o.sizes([len(p["entities"][d]) for d in range(4)])
        for d in range(4):
            for e in p["entities"][d]:
                _entity(o, e, d, True)
        o.text("\n$EndPartitionedEntities\n" if binary else "$EndPartitionedEntities\n")
    o.text("$Nodes\n")
    all_tags = [np.asarray(b[2], dtype=np.int64) for b in sec.node_blocks]
    total = sum(len(t) for t in all_tags)
    lo = min((int(t.min()) for t in all_tags if len(t)), default=0)
    hi = max((int(t.max()) for t in all_tags if len(t)), default=0)
    o.sizes([len(sec.node_blocks), total, lo, hi])
    for dim, ent, tags, coords in sec.node_blocks:
        o.ints([dim, ent, 0])
        o.sizes([len(tags)])
        if binary:
            o.sizes(tags)
            o.floats(coords)
        else:
            for t in tags:
                o.sizes([t])
            for x in np.asarray(coords).reshape(-1, 3):
                o.floats(x)
    o.text("\n$EndNodes\n" if binary else "$EndNodes\n")
    o.text("$Elements\n")
    etags = [np.asarray(b[3], dtype=np.int64) for b in sec.element_blocks]
    total = sum(len(t) for t in etags)
    lo = min((int(t.min()) for t in etags if len(t)), default=0)
    hi = max((int(t.max()) for t in etags if len(t)), default=0)
    o.sizes([len(sec.element_blocks), total, lo, hi])
    for dim, ent, ty, tags, nodes in sec.element_blocks:
        o.ints([dim, ent, ty])
        o.sizes([len(tags)])
        rows = np.concatenate([np.asarray(tags, dtype=np.int64)[:, None],
                               np.asarray(nodes, dtype=np.int64).reshape(len(tags), -1)], axis=1)
        if binary:
            o.sizes(rows)
        else:
            for r in rows:
                o.sizes(r)
    o.text("\n$EndElements\n" if binary else "$EndElements\n")
    if sec.periodic:
        o.text("$Periodic\n")
        o.sizes([len(sec.periodic)])
        for p in sec.periodic:
            o.ints([p["dim"], p["tag"], p["master"]])
            o.sizes([len(p["affine"])])
            o.floats(p["affine"])
            o.sizes([len(p["nodes"])])
            for a, b in p["nodes"]:
                o.sizes([a, b])
        o.text("\n$EndPeriodic\n" if binary else "$EndPeriodic\n")
    return b"".join(o.parts)
