"""A plain-Python reader of ASCII MSH 4.1 sections for the Gmsh tests: an independent reading of
the same files (token by token, node and element rows by line) that gives ``gmsh.Sections`` for
re-encoding, and counts to check the library's parser against."""
from __future__ import annotations

import numpy as np

from gridcl_amd import gmsh


def _entity(tok, dim: int, partitioned: bool) -> dict:
    e = {"tag": int(next(tok))}
    if partitioned:
        e["parent_dim"] = int(next(tok))
        e["parent_tag"] = int(next(tok))
        e["partitions"] = [int(next(tok)) for _ in range(int(next(tok)))]
    e["min"] = [float(next(tok)) for _ in range(3)]
    e["max"] = e["min"] if dim == 0 else [float(next(tok)) for _ in range(3)]
    e["physical"] = [int(next(tok)) for _ in range(int(next(tok)))]
    e["boundary"] = [] if dim == 0 else [int(next(tok)) for _ in range(int(next(tok)))]
    return e


def read_ascii(path) -> gmsh.Sections:
    text = open(path, encoding="utf-8").read()
    sec = gmsh.Sections()
    chunks = text.split("$End")
    for chunk in chunks:
        head = chunk.find("$")
        if head < 0:
            continue
        body = chunk[head:]
        name, _, rest = body.partition("\n")
        name = name.strip()[1:]
        if name == "PhysicalNames":
            lines = rest.strip().splitlines()[1:]
            for ln in lines:
                a, b, nm = ln.split(" ", 2)
                sec.names.append((int(a), int(b), nm.strip().strip('"')))
        elif name == "Entities":
            tok = iter(rest.split())
            counts = [int(next(tok)) for _ in range(4)]
            sec.entities = [[_entity(tok, d, False) for _ in range(counts[d])] for d in range(4)]
        elif name == "PartitionedEntities":
            tok = iter(rest.split())
            p = {"num_partitions": int(next(tok))}
            p["ghosts"] = [(int(next(tok)), int(next(tok))) for _ in range(int(next(tok)))]
            counts = [int(next(tok)) for _ in range(4)]
            p["entities"] = [[_entity(tok, d, True) for _ in range(counts[d])] for d in range(4)]
            sec.partitioned = p
        elif name == "Periodic":
            tok = iter(rest.split())
            for _ in range(int(next(tok))):
                q = {"dim": int(next(tok)), "tag": int(next(tok)), "master": int(next(tok))}
                q["affine"] = [float(next(tok)) for _ in range(int(next(tok)))]
                q["nodes"] = [(int(next(tok)), int(next(tok))) for _ in range(int(next(tok)))]
                sec.periodic.append(q)
        elif name == "Nodes":
            lines = iter(rest.strip().splitlines()[1:])
            for ln in lines:
                dim, ent, param, n = (int(x) for x in ln.split())
                assert param == 0
                tags = np.array([int(next(lines)) for _ in range(n)], dtype=np.int64)
                coords = np.array([[float(x) for x in next(lines).split()] for _ in range(n)],
                                  dtype=np.float64).reshape(n, 3)
                sec.node_blocks.append((dim, ent, tags, coords))
        elif name == "Elements":
            lines = iter(rest.strip().splitlines()[1:])
            for ln in lines:
                dim, ent, ty, n = (int(x) for x in ln.split())
                rows = [[int(x) for x in next(lines).split()] for _ in range(n)]
                arr = np.array(rows, dtype=np.int64).reshape(n, -1)
                sec.element_blocks.append((dim, ent, ty, arr[:, 0].copy(), arr[:, 1:].copy()))
    return sec
