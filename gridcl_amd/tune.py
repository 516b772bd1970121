"""Autotuner of the factor kernels' lane mapping: ``python -m gridcl_amd.tune [--out FILE]``.

For every workload of TUNE_SET (precision, element, point count, mesh size) and output class it
evaluates, with each mapping (one lane per cell storing from registers; one lane per cell with
LDS-staged 16-byte stores; one wavefront per 16-cell tile), X, J and det J in one call (class 0:
more than three components per point) or det J alone plus X alone (class 1: at most three),
timing the mappings interleaved over several rounds with HIP events, keeps the fastest median,
and writes the table that
csrc/tuning.cpp compiles in (csrc/tuning_table.inc), together with a JSON record of every
timing. The reference derives its strategy policy offline from measurements the same way
(tetraloma-bench decide*.rs -> degree-policy.toml).
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAPPINGS = {0: "kMapCell", 1: "kMapCellStaged", 2: "kMapTile"}
OPTS = {0: 8, 1: 16, 2: 32}  # TG_OPT_MAP_CELL, TG_OPT_MAP_CELL_STAGED, TG_OPT_MAP_TILE

# (cell, degree, gdim, rule: ("gauss", degree) | ("tet", 0/1), boxes per direction)
TUNE_SET = [
    (4, 1, 3, ("tet", 0), 48),
    (4, 1, 3, ("gauss", 2), 48),
    (4, 2, 3, ("tet", 1), 48),
    (4, 2, 3, ("gauss", 4), 40),
    (4, 3, 3, ("gauss", 6), 24),
    (5, 1, 3, ("gauss", 2), 80),
    (5, 2, 3, ("gauss", 4), 64),
    (5, 3, 3, ("gauss", 7), 40),
    (6, 1, 3, ("gauss", 4), 64),
    (6, 2, 3, ("gauss", 4), 48),
    (7, 1, 3, ("gauss", 4), 40),
    (2, 3, 3, ("gauss", 4), 384),
    (2, 2, 3, ("gauss", 4), 384),
    (3, 2, 3, ("gauss", 4), 384),
]


def _mesh(M, cell, p, gdim, n):
    if M.TDIM[cell] == 3:
        m = M.cube_box(cell, n, p)
        return M.jitter(m, 0.05 / (n * p), 7, interior_only=True)
    m = M.square_box(cell, n, p, gdim=gdim)
    if gdim == 3:
        m.nodes[:, 2] = 0.1 * np.sin(np.pi * m.nodes[:, 0]) * np.sin(np.pi * m.nodes[:, 1])
    return m


def _time(fn, reps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune(rounds: int = 5, reps: int = 10, quick: bool = False, log=print):
    import torch

    sys.path.insert(0, os.path.dirname(HERE))
    from gridcl_amd import api
    from gridcl_amd import meshes as M

    records = []
    table = []
    for cell, p, gdim, rule, n in TUNE_SET:
        if quick:
            n = max(4, n // 4)
        mesh = _mesh(M, cell, p, gdim, n)
        ps = (api.PointSet.gauss(cell, rule[1]) if rule[0] == "gauss"
              else (api.PointSet.tet_keast4() if rule[1] == 0 else api.PointSet.tet_14()))
        td = M.TDIM[cell]
        classes = ((0, [["points", "jacobians", "detj"]]), (1, [["detj"], ["points"]]))
        k = mesh.blocks[0].nodes_per_cell
        for dtype, tdt, code in (("f32", torch.float32, 0), ("f64", torch.float64, 1)):
            g = api.geometry_from_host(mesh, dtype=tdt)
            C = mesh.blocks[0].num_cells
            out = {nm: torch.empty(int(api._lib.load().tg_field_elems(
                C, len(ps), api.factor_len(nm, td, gdim), 16)), dtype=tdt, device="cuda")
                for nm in ("points", "jacobians", "detj")}
            for cls, calls in classes:
                def run(opt):
                    # The dense kernels' mapping is what is tuned: the sum-factorised
                    # hexahedron kernel (which ignores it) is switched off.
                    for names in calls:
                        api.geometric_factors_block(g, 0, ps, names, options=opt | api.OPT_NO_SUMFACT,
                                                    out={nm: out[nm] for nm in names})
                times = {m: [] for m in MAPPINGS}
                for _ in range(rounds):
                    for m, opt in OPTS.items():
                        times[m].append(_time(lambda: run(opt), reps, 2))
                med = {m: float(np.median(v)) for m, v in times.items()}
                best = min(med, key=med.get)
                rec = {"cell": cell, "degree": p, "gdim": gdim, "tdim": td, "k": k, "nq": len(ps),
                       "dtype": dtype, "class": cls, "cells": C,
                       "ms": {MAPPINGS[m]: med[m] for m in med}, "best": MAPPINGS[best]}
                log(json.dumps(rec))
                records.append(rec)
                table.append((code, td, gdim, k, len(ps), cls, best))
            del g, out
    return records, table


def write_table(table, path: str) -> None:
    """Entries cover each tuned point count exactly; untuned shapes fall back to kMapTile."""
    lines = [
        "// Generated by gridcl_amd/tune.py; edit by re-running it. Entries: {dtype (0 FP32, 1 FP64),",
        "// tdim, gdim, nodes per cell, points low, points high, output class, mapping}.",
        "constexpr uint32_t kDefaultMapping = kMapTile;",
        "const TuneEntry kTuned[] = {",
    ]
    for code, td, gd, k, nq, cls, best in table:
        lines.append(f"    {{{code}, {td}, {gd}, {k}, {nq}, {nq}, {cls}, {MAPPINGS[best]}}},")
    lines.append("};")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "tuning_table.inc"))
    ap.add_argument("--json", default=None)
    ap.add_argument("--rounds", type=int, default=5)
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args()
    records, table = tune(rounds=args.rounds, quick=args.quick,
                          log=lambda s: print(s, flush=True))
    write_table(table, args.out)
    if args.json:
        with open(args.json, "w") as f:
            json.dump(records, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
