"""Benchmark of the geometry hot path on one MI355X.

One step = one fused launch evaluating X, J and det J of every cell of the curved P2
tetrahedral mesh (64^3 hexahedra x 6 Kuhn tetrahedra = 1,572,864 cells, 2,146,689 nodes) at the
14 points of the degree-5 rule, in FP32 (nodes shifted to each cell's first node, FP64 det J at
near-degenerate points). Inputs are resident in HBM before the timed region. The same run times
the same launch on FP64-stored nodes, the plain FP64 kernel on the same mesh, a device copy and a
write-only fill (measured streaming rates), and the C++ port of the reference's CPU path (FP64,
host cores) on the same mesh and on the P1 tetrahedra 32^3 x 6.

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


def profiled_traffic(kernel_sources):
    """HBM bytes per launch of the headline kernel on the 64^3 x 6 mesh, as measured by the two
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, calibrated on a device copy) whose
    summary is committed with the benchmark, and whether that summary is stale: the digest of
    the kernel sources it records differs from the tree's. (None, None) if the summary is
    absent."""
    from gridcl_amd.profsum import source_digest

    traffic = digest = None
    try:
        for line in open(TRAFFIC_PROFILE):
            line = line.strip()
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = int(line.split(":")[1])
            elif line.startswith("kernel_source_digest:"):
                digest = line.split(":")[1].strip()
    except OSError:
        return None, None
    return traffic, digest != source_digest(*kernel_sources)


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


def protocol_seconds(fn):
    """Seconds per call by the reference's timing protocol (tetraloma-bench/src/protocol.rs:1-4,
    58-72; families/degree.rs:1377-1462): one warm-up call, one calibration call of t_cal, then
    ten runs of k = ceil(20 ms / t_cal) calls (1 <= k <= 10,000); the median of the runs' mean
    time per call. Returns (seconds, k)."""
    fn()
    t0 = time.perf_counter()
    fn()
    cal = time.perf_counter() - t0
    calls = int(min(10_000, max(1, np.ceil(0.020 / max(cal, 1e-9)))))
    runs = []
    for _ in range(10):
        t0 = time.perf_counter()
        for _ in range(calls):
            fn()
        runs.append((time.perf_counter() - t0) / calls)
    return float(np.median(runs)), calls


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            return next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")),
                        "unknown")
    except OSError:
        return "unknown"


def host_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))


def cpu_port_leg(mesh, ps, label):
    """The C++ port of the reference's CPU path (oracle/cpu_port.cpp: the E-vector, then the X,
    J and det J launches over tiles of 4 cells held as 4-wide SIMD lanes, node count fixed at
    compile time, multiply-adds fused, OpenMP over tiles) in FP64 on `mesh`'s first block. Each
    launch is timed alone by the reference protocol and the four times are summed, as the
    reference's degree benchmark sums them. Returns the record and the port (outputs filled)."""
    from oracle import oracle as O

    threads = host_threads()
    blk = mesh.blocks[0]
    origin = 0.5 * (mesh.nodes.min(0) + mesh.nodes.max(0))
    values, derivs = O.tabulate(blk.cell, blk.degree, 0, ps.points)
    port = O.CpuPort(blk.nodes_per_cell, mesh.nodes - origin, origin, blk.cells, values, derivs,
                     threads=threads)
    port()
    per, calls = {}, {}
    for name, bit in (("evector", port.EVECTOR), ("x", port.X), ("j", port.J),
                      ("detj", port.DETJ)):
        sec, calls[name] = protocol_seconds(lambda bit=bit: port(bit))
        per[name] = sec * 1e3
    total = sum(per.values())
    evals = blk.num_cells * len(ps)
    return {"value": evals / (total * 1e-3), "unit": "evals/s", "cores": threads, "kind": "port",
            "isa": port.lib.isa, "cpu_model": host_cpu_model(), "host_cpus": os.cpu_count(),
            "ms_per_launch": per, "calls_per_run": calls, "ms_summed": total,
            "sample": label}, port


def untile4(a, cells, nq, length):
    """[tile of 4][q][length][4] -> [cell][q][length]."""
    pad = (cells + 3) // 4 * 4
    a = a.reshape(pad // 4, nq, length, 4).transpose(0, 3, 1, 2)
    return a.reshape(pad, nq, length)[:cells]


def device_rel_err(field, want) -> float:
    """max |field - want| / max |want| over two fields (QuadFields or their tensors) of one
    layout, on the device in FP64."""
    import torch

    a = getattr(field, "data", field).to(torch.float64)
    b = getattr(want, "data", want).to(torch.float64)
    return float((a - b).abs().max() / b.abs().max())


def host_rel_err(field, want) -> float:
    got = field.to_host_cell_major().astype(np.float64)
    return float(np.abs(got - want).max() / np.abs(want).max())


def bench_config1(api, M, steps, warmup, cpu: bool):
    """P1 tetrahedra 32^3 x 6 (196,608 cells, 35,937 nodes) at the 4 Keast points, FP64: the
    GPU kernel (eager and replayed from a graph; a launch moves 86 MB, so eager calls are
    launch-bound) and the CPU port, which must agree to FP64 rounding."""
    import torch

    mesh = M.config1_mesh()
    ps = api.PointSet.tet_keast4()
    blk = mesh.blocks[0]
    C, N, nq = blk.num_cells, mesh.num_nodes, len(ps)
    names = ["points", "jacobians", "detj"]
    g = api.geometry_from_host(mesh, dtype=torch.float64)
    out = {n: torch.empty(int(api._lib.load().tg_field_elems(C, nq, api.factor_len(n, 3, 3), 16)),
                          dtype=torch.float64, device="cuda") for n in names}
    step = lambda: api.geometric_factors_block(g, 0, ps, names, out=out)
    ms = time_steps(step, steps, warmup)
    ms_graph = time_graph(step, steps)
    nbytes = algorithmic_bytes(4, C, N, nq, 8)
    rec = {"cells": C, "nodes": N, "evals": C * nq,
           "gpu_f64": {"ms": ms, "ms_graph": ms_graph, "evals_per_s": C * nq / (ms * 1e-3),
                       "evals_per_s_graph": C * nq / (ms_graph * 1e-3),
                       "achieved_gbs_graph": nbytes / (ms_graph * 1e-3) / 1e9}}
    if cpu:
        cpu_rec, port = cpu_port_leg(mesh, ps, f"P1 tetrahedra 32^3x6 = {C} cells x {nq} Keast "
                                                 "points, whole mesh, FP64")
        fields = {n: api.QuadField(out[n], C, nq, api.factor_len(n, 3, 3), 16) for n in names}
        rec["gpu_f64_vs_cpu_port"] = {
            "points": host_rel_err(fields["points"], untile4(port.x, C, nq, 3)),
            "jacobians": host_rel_err(fields["jacobians"], untile4(port.j, C, nq, 9)),
            "detj": host_rel_err(fields["detj"], untile4(port.detj, C, nq, 1))}
        rec["cpu"] = cpu_rec
    return rec


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


def adjoint_bytes(k: int, cells: int, nodes: int, s: int):
    """Algorithmic bytes of the functional's forward (dofmap 4kC + unique nodes 3sN) and vjp
    (cell pass: dofmap and CSR slots 8kC, nodes 3sN, contributions written 3skC; node pass:
    contributions read 3skC, CSR offsets 8N, FP64 gradient written 24N)."""
    return 4 * k * cells + 3 * s * nodes, 8 * k * cells + 3 * s * nodes + 6 * s * k * cells + 32 * nodes


def bench_adjoint(api, M, mesh, steps, warmup):
    """Forward functional and its vjp on a P2 tetrahedral mesh at the 14-point rule, FP32 and
    FP64: eager calls and graph replays, with each one's fraction of 8 TB/s by the algorithmic
    bytes of adjoint_bytes()."""
    import torch

    h = mesh
    ps = api.PointSet.tet_14()
    blk = h.blocks[0]
    out = {}
    for dtype, tdt, s in (("f32", torch.float32, 4), ("f64", torch.float64, 8)):
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
        bf, bb = adjoint_bytes(blk.nodes_per_cell, blk.num_cells, h.num_nodes, s)
        gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
        out[dtype] = {"ms_forward": ms_f, "ms_vjp": ms_b, "vjp_over_forward": ms_b / ms_f,
                      "ms_forward_graph": ms_fg, "ms_vjp_graph": ms_bg,
                      "vjp_over_forward_graph": ms_bg / ms_fg,
                      "bytes_forward": bf, "bytes_vjp": bb,
                      "roofline_forward": {"achieved_gbs": gbs(bf, ms_fg),
                                           "frac": gbs(bf, ms_fg) / PEAK_HBM_GBS},
                      "roofline_vjp": {"achieved_gbs": gbs(bb, ms_bg),
                                       "frac": gbs(bb, ms_bg) / PEAK_HBM_GBS}}
        del fn, g
    out["cells"] = blk.num_cells
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
    """Q3 hexahedra (96^3 cells, 64 nodes and 64 Gauss points per cell), X + J + det J: the FP32
    sum-factorised kernel with its second and third sums on the matrix cores (the default), the
    same on the vector ALU, FP32 on FP64-stored nodes (matrix cores with FP64 first sums split
    hi/lo, and the vector ALU with every sum in FP64), the dense matrix-core and the dense
    register kernels (sum factorisation off), and the FP64 kernel; each FP32 result's error against the FP64 kernel over the whole mesh."""
    import torch

    h = M.config3_mesh(n=n)
    ps = api.PointSet.gauss(5, 7)
    names = ["points", "jacobians", "detj"]
    C, N = h.blocks[0].num_cells, h.num_nodes
    lib = api._lib.load()

    def buffers(tdt):
        return {nm: torch.empty(int(lib.tg_field_elems(C, 64, api.factor_len(nm, 3, 3), 16)),
                                dtype=tdt, device="cuda") for nm in names}

    g64 = api.geometry_from_host(h, dtype=torch.float64)
    want = buffers(torch.float64)
    step64 = lambda: api.geometric_factors_block(g64, 0, ps, names, out=want)
    ms64 = time_steps(step64, max(2, steps // 2), 1)
    del g64
    g = api.geometry_from_host(h, dtype=torch.float32)
    out = buffers(torch.float32)
    rec = {"cells": C, "nodes": N}
    nbytes = algorithmic_bytes(64, C, N, 64, 4)
    for key, opt in (("mfma", 0), ("valu", api.OPT_NO_MFMA)):
        step = lambda opt=opt: api.geometric_factors_block(g, 0, ps, names, out=out, options=opt)
        spin_up(step, 0.3)
        ms = time_steps(step, steps, warmup)
        torch.cuda.synchronize()
        rec[key] = {"ms": ms, "evals_per_s": C * 64 / (ms * 1e-3),
                    "achieved_gbs": nbytes / (ms * 1e-3) / 1e9,
                    "frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                    "rel_err_vs_fp64_kernel": {nm: device_rel_err(out[nm], want[nm])
                                               for nm in names}}
    rec["ms"], rec["frac"] = rec["mfma"]["ms"], rec["mfma"]["frac"]
    # Sum factorisation off: the dense matrix-core kernel, and the dense register kernel.
    rec["ms_dense_matrix_cores"] = time_steps(lambda: api.geometric_factors_block(
        g, 0, ps, names, out=out, options=api.OPT_NO_SUMFACT), max(2, steps // 5), 1)
    rec["ms_dense_register_kernel"] = time_steps(lambda: api.geometric_factors_block(
        g, 0, ps, names, out=out, options=api.OPT_NO_SUMFACT | api.OPT_NO_MFMA),
        max(2, steps // 5), 1)
    g.set_nodes_f64(torch.tensor(h.nodes - np.asarray(g.origin), dtype=torch.float64,
                                 device="cuda"))
    nb64 = nbytes + 3 * 4 * N
    # FP64 nodes: the matrix-core kernel (FP64 first sums, hi/lo products; the default) and the
    # vector-ALU kernel with every sum in FP64.
    for key, opt in (("fp32_nodes64", 0), ("fp32_nodes64_valu", api.OPT_NO_MFMA)):
        step = lambda opt=opt: api.geometric_factors_block(g, 0, ps, names, out=out, options=opt)
        spin_up(step, 0.3)
        ms = time_steps(step, steps, warmup)
        torch.cuda.synchronize()
        rec[key] = {"ms": ms, "evals_per_s": C * 64 / (ms * 1e-3),
                    "achieved_gbs": nb64 / (ms * 1e-3) / 1e9,
                    "frac": nb64 / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                    "rel_err_vs_fp64_kernel": {nm: device_rel_err(out[nm], want[nm])
                                               for nm in names}}
    rec["fp64"] = {"ms": ms64, "evals_per_s": C * 64 / (ms64 * 1e-3)}
    return rec


def bench_p3tet(api, M, steps, warmup, n=48):
    """Curved P3 tetrahedra (48^3 x 6 = 663,552 cells, 20 nodes per cell), X + J + det J in
    FP32 at the 14-point rule and at Gauss degree 6 (64 points): the dense matrix-core kernel
    (the default) and the register kernel (TG_OPT_NO_MFMA), whose results are equal bit for
    bit, on FP32 nodes and on FP64-stored nodes, and the FP64 kernel; the FP32 results' error
    against the FP64 kernel over the mesh."""
    import torch

    h = M.p3tet_mesh(n=n)
    names = ["points", "jacobians", "detj"]
    C, N = h.blocks[0].num_cells, h.num_nodes
    lib = api._lib.load()
    rec = {"cells": C, "nodes": N}
    g64 = api.geometry_from_host(h, dtype=torch.float64)
    g = api.geometry_from_host(h, dtype=torch.float32)
    gn = api.geometry_from_host(h, dtype=torch.float32, nodes_f64=True)
    for rule, ps in (("tet14", api.PointSet.tet_14()), ("gauss6", api.PointSet.gauss(4, 6))):
        nq = len(ps)

        def buffers(tdt):
            return {nm: torch.empty(int(lib.tg_field_elems(C, nq, api.factor_len(nm, 3, 3), 16)),
                                    dtype=tdt, device="cuda") for nm in names}

        want = buffers(torch.float64)
        ms64 = time_steps(lambda: api.geometric_factors_block(g64, 0, ps, names, out=want),
                          max(2, steps // 2), 1)
        out, reg = buffers(torch.float32), buffers(torch.float32)
        nbytes = algorithmic_bytes(20, C, N, nq, 4)
        r = {"points": nq}
        for key, opt, dst in (("mfma", 0, out), ("valu", api.OPT_NO_MFMA, reg)):
            step = lambda opt=opt, dst=dst: api.geometric_factors_block(g, 0, ps, names, out=dst,
                                                                        options=opt)
            spin_up(step, 0.3)
            ms = time_steps(step, steps, warmup)
            r[key] = {"ms": ms, "evals_per_s": C * nq / (ms * 1e-3),
                      "achieved_gbs": nbytes / (ms * 1e-3) / 1e9,
                      "frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
        torch.cuda.synchronize()
        r["rel_err_vs_fp64_kernel"] = {nm: device_rel_err(out[nm], want[nm]) for nm in names}
        r["mfma_equals_valu"] = all(torch.equal(out[nm], reg[nm]) for nm in names)
        # FP32 on FP64-stored nodes (offsets from node 0 in FP64, rounded once): the matrix
        # cores and the register kernel, against the FP64 kernel on the same mesh.
        nb64 = nbytes + 3 * 4 * N
        for key, opt, dst in (("fp32_nodes64", 0, out), ("fp32_nodes64_valu", api.OPT_NO_MFMA,
                                                          reg)):
            step = lambda opt=opt, dst=dst: api.geometric_factors_block(gn, 0, ps, names,
                                                                        out=dst, options=opt)
            spin_up(step, 0.3)
            ms = time_steps(step, steps, warmup)
            r[key] = {"ms": ms, "evals_per_s": C * nq / (ms * 1e-3),
                      "frac": nb64 / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
        torch.cuda.synchronize()
        r["fp32_nodes64"]["rel_err_vs_fp64_kernel"] = {nm: device_rel_err(out[nm], want[nm])
                                                       for nm in names}
        r["fp32_nodes64"]["equals_valu"] = all(torch.equal(out[nm], reg[nm]) for nm in names)
        r["fp64"] = {"ms": ms64, "evals_per_s": C * nq / (ms64 * 1e-3)}
        rec[rule] = r
        del want, out, reg
    del g64, g, gn
    return rec


def bench_dense_shapes(api, M, steps, warmup):
    """The other 3D cells the dense matrix-core kernel takes (X + J + det J, FP32 and FP64,
    jittered unit boxes of 10,368-524,288 cells): its time per launch against the kernel the
    cell runs on without it (TG_OPT_NO_MFMA; in FP64 the register kernels up to 27 nodes and the
    runtime-K or cooperative kernels above; in FP32 the register kernel of Q2/Q3 hexahedra at a
    point set without tensor structure, the runtime-K kernel of P3/P4 pyramids, P4 tetrahedra,
    P3 prisms and Q4 hexahedra, the cooperative kernel of P5-P8 tetrahedra and P4/P5 prisms),
    and whether its values equal the runtime-K kernel's (TG_OPT_FORCE_GENERIC) bit for bit."""
    import torch

    names = ["points", "jacobians", "detj"]
    hexpts = lambda n: api.PointSet.custom(
        M.HEXAHEDRON, M.splitmix64_uniform(3, 3 * n).reshape(n, 3), np.full(n, 1.0 / n))
    cases = [("q2hex_80", M.HEXAHEDRON, 2, 80, lambda: hexpts(40)),
             ("p3pyr_24", M.PYRAMID, 3, 24, lambda: api.PointSet.gauss(M.PYRAMID, 4)),
             ("p4tet_24", M.TETRAHEDRON, 4, 24, lambda: api.PointSet.gauss(M.TETRAHEDRON, 5)),
             ("p3prism_48", M.PRISM, 3, 48, lambda: api.PointSet.gauss(M.PRISM, 5)),
             ("p5tet_20", M.TETRAHEDRON, 5, 20, lambda: api.PointSet.gauss(M.TETRAHEDRON, 6)),
             ("q3hex_64", M.HEXAHEDRON, 3, 64, lambda: hexpts(100)),
             ("p4prism_32", M.PRISM, 4, 32, lambda: api.PointSet.gauss(M.PRISM, 5)),
             ("p6tet_16", M.TETRAHEDRON, 6, 16, lambda: api.PointSet.gauss(M.TETRAHEDRON, 7)),
             ("p4pyr_16", M.PYRAMID, 4, 16, lambda: api.PointSet.gauss(M.PYRAMID, 5)),
             ("p7tet_14", M.TETRAHEDRON, 7, 14, lambda: api.PointSet.gauss(M.TETRAHEDRON, 8)),
             ("q4hex_32", M.HEXAHEDRON, 4, 32, lambda: hexpts(64)),
             ("p5prism_20", M.PRISM, 5, 20, lambda: api.PointSet.gauss(M.PRISM, 6)),
             ("p8tet_12", M.TETRAHEDRON, 8, 12, lambda: api.PointSet.gauss(M.TETRAHEDRON, 9))]
    out = {}
    for key, cell, p, n, make in cases:
        mesh = M.jitter(M.cube_box(cell, n, p), 0.02 / (n * p), 5, interior_only=True)
        ps = make()
        C, N, k, nq = (mesh.blocks[0].num_cells, mesh.num_nodes, mesh.blocks[0].nodes_per_cell,
                       len(ps))
        out[key] = {"cells": C, "nodes_per_cell": k, "points": nq}
        for dt, tdt, size in (("f32", torch.float32, 4), ("f64", torch.float64, 8)):
            g = api.geometry_from_host(mesh, dtype=tdt)
            a = {nm: f.data for nm, f in api.geometric_factors_block(g, 0, ps, names).items()}
            b = {nm: f.data for nm, f in api.geometric_factors_block(
                g, 0, ps, names, options=api.OPT_FORCE_GENERIC).items()}
            torch.cuda.synchronize()
            rec = {"equals_runtime_k": all(torch.equal(a[nm], b[nm]) for nm in names)}
            nbytes = algorithmic_bytes(k, C, N, nq, size)
            for label, opt, dst in (("mfma", 0, a), ("without", api.OPT_NO_MFMA, b)):
                step = lambda opt=opt, dst=dst: api.geometric_factors_block(
                    g, 0, ps, names, out=dst, options=opt)
                spin_up(step, 0.2)
                ms = time_steps(step, steps, warmup)
                rec[label] = {"ms": ms, "frac": nbytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
            rec["speedup"] = rec["without"]["ms"] / rec["mfma"]["ms"]
            out[key][dt] = rec
            del g, a, b
            torch.cuda.empty_cache()
    return out


HEADLINE_SOURCES = ("factor_kernels.hip",)
EXTRAS = ("config1", "adjoint", "hybrid", "q3", "p3tet", "dense", "pullback", "measure",
          "equispaced", "gmsh")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=64, help="hexahedra per direction (64: full size)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spin-up", type=float, default=0.5,
                    help="seconds of untimed back-to-back steps before the warm-up")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the config-1, Q3, mixed-mesh, adjoint and other legs")
    ap.add_argument("--extras", default=",".join(EXTRAS),
                    help="comma-separated extra legs to run (default: all of "
                         f"{','.join(EXTRAS)})")
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

    # (key, output type, node storage): FP32 on FP32 nodes, FP32 on FP64 nodes, FP64.
    results = {}
    fields = {}
    for key, tdt, s_out, nodes64 in (("f32", torch.float32, 4, False),
                                     ("f32_nodes64", torch.float32, 4, True),
                                     ("f64", torch.float64, 8, False)):
        g = api.geometry_from_host(mesh, dtype=tdt, nodes_f64=nodes64)
        out = {n: torch.empty(int(api._lib.load().tg_field_elems(C, nq, api.factor_len(n, 3, 3),
                                                                  16)),
                              dtype=tdt, device="cuda") for n in names}
        step = lambda g=g, out=out: api.geometric_factors_block(g, 0, ps, names, tile=16,
                                                                out=out)
        spun = spin_up(step, args.spin_up)
        ms = time_steps(step, args.steps, args.warmup)
        nbytes = algorithmic_bytes(k, C, N, nq, s_out) + (3 * 4 * N if nodes64 else 0)
        results[key] = {"ms": ms, "evals_per_s": C * nq / (ms * 1e-3), "bytes": nbytes,
                        "achieved_gbs": nbytes / (ms * 1e-3) / 1e9}
        fields[key] = {n: api.QuadField(out[n], C, nq, api.factor_len(n, 3, 3), 16)
                       for n in names}
        log(f"{key}: {spun} spin-up calls, {ms * 1e3:.1f} us/step, "
            f"{results[key]['evals_per_s'] / 1e9:.1f} G evals/s, "
            f"{results[key]['achieved_gbs']:.0f} GB/s")
        del g
    ceiling, write_ceiling = stream_ceilings_gbs()
    log(f"device copy {ceiling:.0f} GB/s, fill (write only) {write_ceiling:.0f} GB/s")

    # FP32 against the FP64 kernel over the whole mesh: with FP32 node storage the rounding of
    # the stored coordinates dominates; with FP64 node storage only the FP32 arithmetic is left.
    vs_fp64 = {key: {n: device_rel_err(fields[key][n].data, fields["f64"][n].data)
                     for n in names} for key in ("f32", "f32_nodes64")}
    log(f"FP32 vs FP64 kernel (whole mesh): {vs_fp64}")

    cpu = None
    if not args.no_cpu:
        cpu, port = cpu_port_leg(mesh, ps, f"curved P2 tetrahedra {args.n}^3x6 = {C} cells x "
                                           f"{nq} points, whole mesh, FP64")
        log(f"cpu port {cpu['value'] / 1e9:.3f} G evals/s on {cpu['cores']} threads "
            f"({cpu['isa']}), per launch ms {cpu['ms_per_launch']}")
        # The CPU port and the GPU FP64 kernel compute the same FP64 values (to rounding).
        cpu["gpu_f64_vs_port"] = {
            "points": host_rel_err(fields["f64"]["points"], untile4(port.x, C, nq, 3)),
            "jacobians": host_rel_err(fields["f64"]["jacobians"], untile4(port.j, C, nq, 9)),
            "detj": host_rel_err(fields["f64"]["detj"], untile4(port.detj, C, nq, 1))}
        log(f"GPU FP64 vs CPU port: {cpu['gpu_f64_vs_port']}")
        del port

    extras = None
    legs = set() if args.no_extras else {x for x in args.extras.split(",") if x}
    unknown = legs - set(EXTRAS)
    if unknown:
        print(f"unknown extra legs: {sorted(unknown)}", file=sys.stderr)
        return 2
    if legs:
        del fields
        torch.cuda.empty_cache()
        extras = {}
    if "config1" in legs:
        c1 = bench_config1(api, M, args.steps, args.warmup, not args.no_cpu)
        if cpu is not None:
            cpu["config1"] = c1.pop("cpu")
            cpu["config1"]["gpu_f64_vs_port"] = c1.pop("gpu_f64_vs_cpu_port")
        extras["config1_p1tet_32_f64"] = c1
        log(f"config 1: {c1}; cpu {cpu and cpu.get('config1')}")
    if "adjoint" in legs:
        extras["adjoint_p2tet_32"] = bench_adjoint(api, M, M.config5_mesh(), args.steps,
                                                   args.warmup)
        log(f"adjoint 32: {extras['adjoint_p2tet_32']}")
        extras["adjoint_p2tet_64"] = bench_adjoint(api, M, mesh, args.steps, args.warmup)
        log(f"adjoint 64: {extras['adjoint_p2tet_64']}")
        torch.cuda.empty_cache()
    if "hybrid" in legs:
        extras["hybrid"] = bench_hybrid(api, M, args.steps, args.warmup)
        log(f"hybrid: {extras['hybrid']}")
        torch.cuda.empty_cache()
    if "q3" in legs:
        extras["q3_hex_96"] = bench_q3(api, M, max(5, args.steps // 5), 2)
        log(f"q3: {extras['q3_hex_96']}")
        torch.cuda.empty_cache()
    if "p3tet" in legs:
        extras["p3tet_48"] = bench_p3tet(api, M, args.steps, args.warmup)
        log(f"p3 tetrahedra: {extras['p3tet_48']}")
        torch.cuda.empty_cache()
    if "dense" in legs:
        extras["dense_shapes"] = bench_dense_shapes(api, M, max(5, args.steps // 5), 2)
        log(f"dense shapes: {extras['dense_shapes']}")
        torch.cuda.empty_cache()
    if "pullback" in legs:
        extras["pull_back_p2tet_64"] = bench_pull_back(api, M, args.steps, args.warmup)
        log(f"pull-back: {extras['pull_back_p2tet_64']}")
    if "measure" in legs:
        extras["measure_p2tet_48"] = bench_measure(api, M, args.steps, args.warmup)
        log(f"measure: {extras['measure_p2tet_48']}")
    if "equispaced" in legs:
        extras["equispaced_q3hex_64"] = bench_equispaced(api, M, args.steps, args.warmup)
        log(f"equispaced: {extras['equispaced_q3hex_64']}")
        torch.cuda.empty_cache()
    if "gmsh" in legs:
        extras["gmsh_read"] = bench_gmsh(api, M, mesh)
        log(f"gmsh: {extras['gmsh_read']}")

    r32, r64, rn = results["f32"], results["f64"], results["f32_nodes64"]
    traffic, stale = profiled_traffic(HEADLINE_SOURCES) if args.n == 64 else (None, None)
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
                     "traffic": traffic, "traffic_profile_stale": stale},
        "cpu_baseline": cpu,
        "fp32_nodes64": {"evals_per_s": rn["evals_per_s"], "ms_per_step": rn["ms"],
                         "achieved_gbs": rn["achieved_gbs"],
                         "frac": rn["achieved_gbs"] / PEAK_HBM_GBS,
                         "rel_err_vs_fp64_kernel": vs_fp64["f32_nodes64"]},
        "fp64": {"evals_per_s": r64["evals_per_s"], "ms_per_step": r64["ms"],
                 "achieved_gbs": r64["achieved_gbs"],
                 "frac": r64["achieved_gbs"] / PEAK_HBM_GBS},
        "fp32_over_fp64": r32["evals_per_s"] / r64["evals_per_s"],
        "spin_up_s": args.spin_up,
        "copy_gbs": ceiling,
        "write_ceiling_gbs": write_ceiling,
        "frac_of_write_ceiling": r32["achieved_gbs"] / write_ceiling,
        "fp32_rel_err_vs_fp64_kernel": vs_fp64["f32"],
        "extras": extras,
    }
    if extras and "adjoint_p2tet_32" in extras:
        line["roofline_adjoint"] = {
            name: {dt: {"forward": extras[name][dt]["roofline_forward"]["frac"],
                        "vjp": extras[name][dt]["roofline_vjp"]["frac"]}
                   for dt in ("f32", "f64")}
            for name in ("adjoint_p2tet_32", "adjoint_p2tet_64")}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
