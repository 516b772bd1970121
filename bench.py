"""Benchmark of the geometry hot path on one MI355X.

One step = one fused launch evaluating X, J and det J of every cell of the curved P2
tetrahedral mesh (64^3 hexahedra x 6 Kuhn tetrahedra = 1,572,864 cells, 2,146,689 nodes) at the
14 points of the degree-5 rule, in FP32 (nodes shifted to each cell's first node, FP64 det J at
near-degenerate points). Inputs are resident in HBM before the timed region. The same run times
the plain FP64 kernel on the same mesh, a device copy and a write-only fill (measured streaming
rates), and the FP64 CPU port of the reference's CPU path on a bounded sample of the same mesh.

Prints one JSON line. ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X published HBM3E bandwidth
METRIC = "cell×qpt geometry evals/s + % HBM roofline (fwd, adjoint), FP32 vs FP64"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(k: int, cells: int, nodes: int, nq: int, s: int, length: int = 13) -> int:
    """dofmap 4kC + unique node coordinates 3sN + outputs LEN s nq C (X 3 + J 9 + det J 1)."""
    return 4 * k * cells + 3 * s * nodes + length * s * nq * cells


def time_steps(fn, steps: int, warmup: int):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps  # ms per step


def time_graph(fn, replays: int, calls: int = 20):
    """ms per call of `fn`, from `replays` replays of a graph that holds `calls` calls."""
    import torch

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(calls):
            fn()
    graph.replay()
    torch.cuda.synchronize()
    ms = time_steps(graph.replay, max(1, replays), 1)
    del graph
    return ms / calls


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "factor_direct_p2tet_64cube_hbm_traffic.txt")


def profiled_traffic():
    """HBM bytes per launch of the FP32 kernel on the 64^3 x 6 mesh, as measured by the two
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, calibrated on a device copy) whose
    summary is committed with the benchmark; None if that summary is absent."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def spin_up(fn, seconds: float) -> int:
    """Runs `fn` back to back for about `seconds` of wall time, untimed, so that the GPU clocks
    have left their idle state before the warm-up and timed steps. Returns the number of calls."""
    import torch

    calls = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            fn()
        torch.cuda.synchronize()
        calls += 20
    return calls


def stream_ceilings_gbs() -> tuple[float, float]:
    """Measured streaming rates on 2 GiB: torch's device copy (read + write bytes over time) and
    torch's fill (write only; the benchmark kernels move 93 % of their bytes as writes)."""
    import torch

    n = 1 << 29
    a = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    ms_copy = time_steps(lambda: b.copy_(a), 10, 3)
    ms_fill = time_steps(lambda: b.fill_(2.0), 10, 3)
    del a, b
    return 2 * n * 4 / (ms_copy * 1e-3) / 1e9, n * 4 / (ms_fill * 1e-3) / 1e9


def cpu_baseline(mesh, ps, sample_cells: int):
    """The oracle's CPU path (E-vector gather, then X, J, det J in f64 into tiles of 4, OpenMP
    on the host cores) on the first `sample_cells` cells, timed with the reference protocol
    (warm-up, calibration to >= 1 s per run, 10 runs, median)."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    blk = mesh.blocks[0]
    cells = np.ascontiguousarray(blk.cells[:sample_cells])
    # The coordinates the FP32 geometry stores (relative to the bounding-box centre, rounded to
    # FP32), widened to FP64: the same mesh the GPU evaluates.
    origin = 0.5 * (mesh.nodes.min(0) + mesh.nodes.max(0))
    rel = (mesh.nodes - origin).astype(np.float32).astype(np.float64)
    values, derivs = O.tabulate(blk.cell, blk.degree, 0, ps.points)
    run = lambda: O.cpu_factors(3, 3, blk.nodes_per_cell, rel, origin, cells, values, derivs,
                                tile=4, threads=threads)
    out = run()
    t0 = time.perf_counter()
    run()
    t1 = time.perf_counter() - t0
    calls = max(1, int(np.ceil(1.0 / max(t1, 1e-6))))
    times = []
    for _ in range(10):
        t0 = time.perf_counter()
        for _ in range(calls):
            run()
        times.append((time.perf_counter() - t0) / calls)
    med = float(np.median(times))
    evals = sample_cells * len(ps)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                         model)
    except OSError:
        pass
    return {"value": evals / med, "unit": "evals/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "sample": f"first {sample_cells} cells of the benchmark mesh x {len(ps)} points "
                      f"(FP64, E-vector + X + J + det J, tiles of 4), median of 10 runs of "
                      f"{calls} calls"}, out


def bench_hybrid(api, M, steps, warmup, scale=1.0):
    """The mixed mesh (P2 tets, Q2 hexes, P2 prisms, P1 pyramids, P3 surface triangles with
    normals; Gauss degree 4): the persistent launch and its graph replay, per-block launches, and
    the per-block plan replayed from one graph."""
    import torch

    h = M.config4_mesh(scale=scale)
    g = api.geometry_from_host(h, dtype=torch.float32)
    rule = api.QuadRule.for_grid(g, 4)
    factors = ["points", "jacobians", "detj", "normals"]
    plan = api.MeshPlan(g, rule, factors, options=api.OPT_PLAN_PERSISTENT)
    evals = sum(b.num_cells * len(rule.get(b.cell)) for b in h.blocks)
    nbytes = 0
    for b, blk in enumerate(h.blocks):
        nq = len(rule.get(blk.cell))
        td = M.TDIM[blk.cell]
        length = 3 + 3 * td + 1 + (3 if td == 2 else 0)
        nbytes += 4 * blk.nodes_per_cell * blk.num_cells + length * 4 * nq * blk.num_cells
    nbytes += 3 * 4 * h.num_nodes
    plan_items, plan_grid = plan.num_items, plan.grid
    assert plan.persistent
    ms_plan = time_steps(plan.launch, steps, warmup)
    ms_graph = time_steps(plan.graph_launch, steps, warmup)
    outs = [{n: f.data for n, f in fb.items()} for fb in plan.fields]
    sets = [rule.get(b.cell) for b in h.blocks]

    def per_block():
        for b in range(len(h.blocks)):
            api.geometric_factors_block(g, b, sets[b], list(outs[b]), out=outs[b])

    ms_blocks = time_steps(per_block, steps, warmup)
    del plan
    pb = api.MeshPlan(g, rule, factors, options=api.OPT_PLAN_PER_BLOCK)
    ms_pb_graph = time_steps(pb.graph_launch, steps, warmup)
    best = min(ms_graph, ms_pb_graph)
    return {
        "blocks": [(int(b.cell), int(b.degree), b.num_cells) for b in h.blocks],
        "evals": evals, "work_items": plan_items, "grid": plan_grid,
        "ms_persistent": ms_plan, "ms_persistent_graph": ms_graph,
        "ms_per_block_launches": ms_blocks, "ms_per_block_graph": ms_pb_graph,
        "evals_per_s_best_graph": evals / (best * 1e-3),
        "achieved_gbs_best_graph": nbytes / (best * 1e-3) / 1e9,
    }


def bench_adjoint(api, M, steps, warmup):
    """Forward functional and its vjp on the config-5 mesh (P2 tets 32^3 x 6, 14 points)."""
    import torch

    h = M.config5_mesh()
    ps = api.PointSet.tet_14()
    out = {}
    for dtype, tdt in (("f32", torch.float32), ("f64", torch.float64)):
        g = api.geometry_from_host(h, dtype=tdt)
        fn = api.MeshFunctional(g, 0, ps, beta=0.1)
        val = torch.empty((), dtype=torch.float64, device="cuda")
        grad = torch.empty(g.nodes.shape, dtype=torch.float64, device="cuda")
        lib = api._lib.load()

        def fwd():
            lib.tg_functional_forward(fn._h, None, 0.1, val.data_ptr(), api._stream())

        def vjp():
            lib.tg_functional_vjp(fn._h, None, 0.1, None, grad.data_ptr(), api._stream())

        ms_f = time_steps(fwd, steps, warmup)
        ms_b = time_steps(vjp, steps, warmup)
        vjp()
        torch.cuda.synchronize()
        grad_eager = grad.clone()
        # The same calls replayed from captured graphs (20 calls each): device time only, without
        # the host's launch overhead, which at ~20-60 us per call is of the same order.
        ms_fg = time_graph(fwd, steps)
        ms_bg = time_graph(vjp, steps)
        torch.cuda.synchronize()
        assert torch.equal(grad, grad_eager), "graph-replayed vjp differs from the eager one"
        out[dtype] = {"ms_forward": ms_f, "ms_vjp": ms_b, "vjp_over_forward": ms_b / ms_f,
                      "ms_forward_graph": ms_fg, "ms_vjp_graph": ms_bg,
                      "vjp_over_forward_graph": ms_bg / ms_fg}
        del fn, g
    out["cells"] = h.blocks[0].num_cells
    out["nodes"] = h.num_nodes
    return out


def bench_pull_back(api, M, steps, warmup, n=64):
    """Pull-back of the 14 Gauss points of every cell of the curved P2 tetrahedra 64^3 x 6
    (the headline mesh), 16 Newton steps per point (the reference's default), FP32 and FP64:
    the physical points are computed once, then pulled back; the round-trip error is reported."""
    import torch

    h = M.config2_mesh(n=n)
    ps = api.PointSet.tet_14()
    out = {"cells": h.blocks[0].num_cells, "points_per_cell": len(ps), "iterations": 16}
    for dtype, tdt in (("f32", torch.float32), ("f64", torch.float64)):
        g = api.geometry_from_host(h, dtype=tdt)
        x = api.physical_points_block(g, 0, ps)
        step = lambda: api.pull_back_block(g, 0, x)
        ms = time_steps(step, max(2, steps // 5), 1)
        xi, last = api.pull_back_block(g, 0, x)
        torch.cuda.synchronize()
        # Tiles of 16 cells: [tile][q][3][16]; every cell's points should come back.
        want = torch.tensor(ps.points, dtype=tdt, device="cuda")[None, :, :, None]
        err = float((xi.data.view(-1, len(ps), 3, 16) - want).abs().max())
        pts = h.blocks[0].num_cells * len(ps)
        out[dtype] = {"ms": ms, "points_per_s": pts / (ms * 1e-3), "max_xi_error": err,
                      "max_last_step": float(last.data.max())}
        del g, x, xi, last
        torch.cuda.empty_cache()
    return out


def gmsh_sections(M, mesh):
    """The sections of an MSH 4.1 file holding `mesh`'s one cell block: node tags 1..N in one
    block, element tags 1..C, each cell's nodes in Gmsh's order."""
    from gridcl_amd import gmsh

    blk = mesh.blocks[0]
    order = gmsh.gmsh_to_internal(blk.cell, blk.degree)
    cells = blk.cells.reshape(blk.num_cells, -1)
    coords = np.zeros((mesh.num_nodes, 3))
    coords[:, :mesh.gdim] = mesh.nodes
    ty = {(4, 2): 11, (4, 1): 4}[(blk.cell, blk.degree)]
    return gmsh.Sections(
        node_blocks=[(3, 1, np.arange(1, mesh.num_nodes + 1, dtype=np.int64), coords)],
        element_blocks=[(3, 1, ty, np.arange(1, blk.num_cells + 1, dtype=np.int64),
                         cells[:, order].astype(np.int64) + 1)])


def bench_gmsh(api, M, mesh, runs=5):
    """Reading the headline mesh (curved P2 tetrahedra 64^3 x 6) from a binary MSH 4.1 file:
    the host reader (parse, unique-tag checks, Gmsh-to-A.5 permutation, u32 dofmap) from memory
    and from a file in the page cache, then the upload into a device geometry; the P2 mesh
    32^3 x 6 as an ASCII file. The dofmap read back must equal the generator's bit for bit."""
    import ctypes
    import tempfile

    import torch

    from gridcl_amd import _lib, gmsh

    lib = _lib.load()
    out = {}
    for name, m, binary in (("binary_p2tet_64", mesh, True),
                            ("ascii_p2tet_32", M.config5_mesh(), False)):
        data = gmsh.encode(gmsh_sections(M, m), binary)
        times = []
        for _ in range(runs):
            h = ctypes.c_void_p()
            t0 = time.perf_counter()
            rc = lib.tg_gmsh_read_bytes(data, len(data), 0, ctypes.byref(h))
            times.append(time.perf_counter() - t0)
            assert rc == 0, _lib.last_error()
            lib.tg_gmsh_destroy(h)
        mem = float(np.median(times))
        with tempfile.NamedTemporaryFile(suffix=".msh") as f:
            f.write(data)
            f.flush()
            times = []
            for _ in range(runs):
                h = ctypes.c_void_p()
                t0 = time.perf_counter()
                rc = lib.tg_gmsh_read_file(f.name.encode(), 0, ctypes.byref(h))
                times.append(time.perf_counter() - t0)
                assert rc == 0, _lib.last_error()
                lib.tg_gmsh_destroy(h)
            disk = float(np.median(times))
            read = gmsh.read(f.name)
        cells = m.blocks[0].cells.reshape(m.blocks[0].num_cells, -1)
        assert np.array_equal(read.blocks[0].nodes, cells)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g = read.geometry(torch.float32)
        torch.cuda.synchronize()
        upload = time.perf_counter() - t0
        del g
        out[name] = {"file_mb": len(data) / 1e6, "cells": int(cells.shape[0]),
                     "nodes": int(m.num_nodes), "read_bytes_s": mem, "read_file_s": disk,
                     "mb_per_s": len(data) / 1e6 / mem, "cells_per_s": cells.shape[0] / mem,
                     "upload_f32_s": upload, "dofmap_identical": True}
    # VTU output of the headline geometry (FP32): download, VTK order, raw appended data.
    from gridcl_amd import vtk

    g = api.geometry_from_host(mesh, dtype=torch.float32)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mesh.vtu")
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            vtk.write_geometry(path, g)
            times.append(time.perf_counter() - t0)
        size = os.path.getsize(path)
    del g
    out["vtu_write_p2tet_64"] = {"file_mb": size / 1e6, "write_s": float(np.median(times)),
                                 "mb_per_s": size / 1e6 / float(np.median(times))}
    return out


def bench_measure(api, M, steps, warmup, n=48):
    """Block measure (sum_c sum_q w_q det J, the reference's fixed-order sums widened to FP64) of
    the curved P2 tetrahedra n^3 x 6 (48: another grid than the headline's, so that profiles keep
    the headline's dispatches apart) at the 14-point rule, FP32 and FP64, on preallocated
    outputs."""
    import torch

    mesh = M.config2_mesh(n=n)
    ps = api.PointSet.tet_14()
    lib = api._lib.load()
    out = {"cells": mesh.blocks[0].num_cells, "points_per_cell": len(ps)}
    for dtype, tdt in (("f32", torch.float32), ("f64", torch.float64)):
        g = api.geometry_from_host(mesh, dtype=tdt)
        val = torch.empty((), dtype=torch.float64, device="cuda")
        call = lambda: lib.tg_measure_block(g.handle, 0, ps._h, 0, val.data_ptr(), api._stream())
        ms = time_steps(call, steps, warmup)
        torch.cuda.synchronize()
        out[dtype] = {"ms": ms, "evals_per_s": mesh.blocks[0].num_cells * len(ps) / (ms * 1e-3),
                      "volume": float(val.item())}
        del g
        torch.cuda.empty_cache()
    return out


def bench_equispaced(api, M, steps, warmup, n=64):
    """GLL -> equispaced node conversion of Q3 hexahedra n^3 read as GLL nodes, FP32: X at each
    cell's 64 equispaced lattice points, then each node's first occurrence."""
    import torch

    h = M.config3_mesh(n=n)
    h = M.HostMesh(h.gdim, h.nodes, [M.HostBlock(b.cell, b.degree, b.cells, family=1)
                                     for b in h.blocks])
    g = api.geometry_from_host(h, dtype=torch.float32)
    lib = api._lib.load()
    out_t = torch.empty_like(g.nodes)
    call = lambda: lib.tg_equispaced_nodes(g.handle, out_t.data_ptr(), api._stream())
    ms = time_steps(call, max(2, steps // 5), 2)
    return {"cells": h.blocks[0].num_cells, "nodes": h.num_nodes, "ms": ms,
            "nodes_per_s": h.num_nodes / (ms * 1e-3)}


def bench_q3(api, M, steps, warmup, n=96):
    """Q3 hexahedra (96^3 cells, 64 nodes and 64 Gauss points per cell), X + J + det J, FP32."""
    import torch

    h = M.config3_mesh(n=n)
    g = api.geometry_from_host(h, dtype=torch.float32)
    ps = api.PointSet.gauss(5, 7)
    names = ["points", "jacobians", "detj"]
    C, N = h.blocks[0].num_cells, h.num_nodes
    out = {nm: torch.empty(int(api._lib.load().tg_field_elems(C, 64, api.factor_len(nm, 3, 3),
                                                               16)),
                           dtype=torch.float32, device="cuda") for nm in names}
    step = lambda: api.geometric_factors_block(g, 0, ps, names, out=out)
    spin_up(step, 0.3)
    ms = time_steps(step, steps, warmup)
    ms_dense = time_steps(lambda: api.geometric_factors_block(
        g, 0, ps, names, out=out, options=api.OPT_NO_SUMFACT), max(2, steps // 5), 1)
    nbytes = algorithmic_bytes(64, C, N, 64, 4)
    return {"cells": C, "nodes": N, "ms": ms, "evals_per_s": C * 64 / (ms * 1e-3),
            "achieved_gbs": nbytes / (ms * 1e-3) / 1e9, "frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "kernel": "sum-factorised FP32 (vertex-mean shift)",
            "ms_dense_register_kernel": ms_dense}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=64, help="hexahedra per direction (64: full size)")
    ap.add_argument("--cpu-sample", type=int, default=262_144)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spin-up", type=float, default=0.5,
                    help="seconds of untimed back-to-back steps before the warm-up")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the Q3, mixed-mesh and adjoint legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): cell blocks would shard "
              "across GPUs by cell ranges, which this benchmark does not run.", file=sys.stderr)
        return 2

    import torch

    from gridcl_amd import api
    from gridcl_amd import meshes as M

    t0 = time.time()
    mesh = M.config2_mesh(n=args.n)
    blk = mesh.blocks[0]
    C, N, k = blk.num_cells, mesh.num_nodes, blk.nodes_per_cell
    log(f"mesh {C} cells {N} nodes in {time.time() - t0:.1f}s")
    ps = api.PointSet.tet_14()
    nq = len(ps)
    names = ["points", "jacobians", "detj"]

    results = {}
    fields = {}
    for dtype, tdt, s in (("f32", torch.float32, 4), ("f64", torch.float64, 8)):
        g = api.geometry_from_host(mesh, dtype=tdt)
        out = {n: torch.empty(int(api._lib.load().tg_field_elems(C, nq, api.factor_len(n, 3, 3),
                                                                  16)),
                              dtype=tdt, device="cuda") for n in names}
        step = lambda g=g, out=out: api.geometric_factors_block(g, 0, ps, names, tile=16,
                                                                out=out)
        spun = spin_up(step, args.spin_up)
        ms = time_steps(step, args.steps, args.warmup)
        nbytes = algorithmic_bytes(k, C, N, nq, s)
        results[dtype] = {"ms": ms, "evals_per_s": C * nq / (ms * 1e-3), "bytes": nbytes,
                          "achieved_gbs": nbytes / (ms * 1e-3) / 1e9}
        fields[dtype] = {n: api.QuadField(out[n], C, nq, api.factor_len(n, 3, 3), 16)
                         for n in names}
        log(f"{dtype}: {spun} spin-up calls, {ms * 1e3:.1f} us/step, {results[dtype]['evals_per_s'] / 1e9:.1f} G evals/s, "
            f"{results[dtype]['achieved_gbs']:.0f} GB/s")
        del g
    ceiling, write_ceiling = stream_ceilings_gbs()
    log(f"device copy {ceiling:.0f} GB/s, fill (write only) {write_ceiling:.0f} GB/s")

    cpu = None
    accuracy = None
    if not args.no_cpu:
        sample = min(args.cpu_sample, C)
        cpu, (x, j, dj) = cpu_baseline(mesh, ps, sample)
        log(f"cpu baseline {cpu['value'] / 1e9:.3f} G evals/s on {cpu['cores']} threads")
        # The CPU leg's FP64 values on its sample against the GPU FP32 fields (rel_err).
        pad = (sample + 3) // 4 * 4

        def untile(a, length, tile=4):
            a = a.reshape(pad // tile, nq, length, tile).transpose(0, 3, 1, 2)
            return a.reshape(pad, nq, length)[:sample]

        want = {"points": untile(x, 3), "jacobians": untile(j, 9), "detj": untile(dj, 1)}
        accuracy = {}
        for n in names:
            got = fields["f32"][n].to_host_cell_major()[:sample].astype(np.float64)
            accuracy[n] = float(np.abs(got - want[n]).max() / np.abs(want[n]).max())
        log(f"FP32 max rel_err vs FP64 on the sample: {accuracy}")

    # FP32 against the FP64 kernel over the whole mesh: rounding of the stored coordinates plus
    # the FP32 arithmetic.
    vs_fp64_kernel = {}
    for n in names:
        a = fields["f32"][n].to_host_cell_major().astype(np.float64)
        b = fields["f64"][n].to_host_cell_major()
        vs_fp64_kernel[n] = float(np.abs(a - b).max() / np.abs(b).max())
        del a, b
    log(f"FP32 vs FP64 kernel (whole mesh, incl. FP32 storage of nodes): {vs_fp64_kernel}")

    extras = None
    if not args.no_extras:
        del fields
        torch.cuda.empty_cache()
        extras = {}
        extras["adjoint_p2tet_32"] = bench_adjoint(api, M, args.steps, args.warmup)
        log(f"adjoint: {extras['adjoint_p2tet_32']}")
        extras["hybrid"] = bench_hybrid(api, M, args.steps, args.warmup)
        log(f"hybrid: {extras['hybrid']}")
        torch.cuda.empty_cache()
        extras["q3_hex_96"] = bench_q3(api, M, max(5, args.steps // 5), 2)
        log(f"q3: {extras['q3_hex_96']}")
        torch.cuda.empty_cache()
        extras["pull_back_p2tet_64"] = bench_pull_back(api, M, args.steps, args.warmup)
        log(f"pull-back: {extras['pull_back_p2tet_64']}")
        extras["measure_p2tet_48"] = bench_measure(api, M, args.steps, args.warmup)
        log(f"measure: {extras['measure_p2tet_48']}")
        extras["equispaced_q3hex_64"] = bench_equispaced(api, M, args.steps, args.warmup)
        log(f"equispaced: {extras['equispaced_q3hex_64']}")
        torch.cuda.empty_cache()
        extras["gmsh_read"] = bench_gmsh(api, M, mesh)
        log(f"gmsh: {extras['gmsh_read']}")

    r32, r64 = results["f32"], results["f64"]
    line = {
        "metric": METRIC,
        "value": r32["evals_per_s"],
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"curved P2 tetrahedra {args.n}^3x6 = {C} cells, {N} nodes, 14-point "
                        "degree-5 rule, X+J+det J in one fused FP32 launch",
            "cells": C, "nodes": N, "points_per_cell": nq, "nodes_per_cell": k, "tile": 16,
        },
        "roofline": {"bound": "hbm", "achieved": r32["achieved_gbs"], "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": r32["achieved_gbs"] / PEAK_HBM_GBS,
                     "traffic": profiled_traffic() if args.n == 64 else None},
        "cpu_baseline": cpu,
        "fp64": {"evals_per_s": r64["evals_per_s"], "ms_per_step": r64["ms"],
                 "achieved_gbs": r64["achieved_gbs"],
                 "frac": r64["achieved_gbs"] / PEAK_HBM_GBS},
        "fp32_over_fp64": r32["evals_per_s"] / r64["evals_per_s"],
        "spin_up_s": args.spin_up,
        "copy_gbs": ceiling,
        "write_ceiling_gbs": write_ceiling,
        "frac_of_write_ceiling": r32["achieved_gbs"] / write_ceiling,
        "fp32_rel_err_vs_fp64_sample": accuracy,
        "fp32_rel_err_vs_fp64_kernel": vs_fp64_kernel,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
