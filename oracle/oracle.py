"""ctypes binding of the FP64 CPU oracle (oracle.c). TEST INFRASTRUCTURE ONLY: imported by
tests/, the smoke check and the CPU-baseline leg of bench.py, never by gridcl_amd/."""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_long, c_size_t, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.or_tdim.restype = c_int
        lib.or_tdim.argtypes = [c_int]
        lib.or_sub_entity.restype = c_int
        lib.or_sub_entity.argtypes = [c_int, c_int, c_int, POINTER(c_int), POINTER(c_int)]
        lib.or_num_nodes.restype = c_size_t
        lib.or_num_nodes.argtypes = [c_int, c_int]
        lib.or_lagrange_points.restype = c_long
        lib.or_lagrange_points.argtypes = [c_int, c_int, c_int, POINTER(c_double)]
        lib.or_tabulate.restype = c_int
        lib.or_tabulate.argtypes = [c_int, c_int, c_int, POINTER(c_double), c_size_t, c_void_p,
                                    c_void_p]
        lib.or_eval_block.restype = c_int
        lib.or_eval_block.argtypes = [c_int, c_int, c_int, c_int, POINTER(c_double), c_size_t,
                                      c_void_p, c_size_t, c_void_p, POINTER(c_double), c_size_t,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_int]
        lib.or_cpu_factors.restype = c_int
        lib.or_cpu_factors.argtypes = [c_int, c_int, c_int, POINTER(c_double), c_size_t,
                                       POINTER(c_double), c_void_p, c_size_t, POINTER(c_double),
                                       POINTER(c_double), c_size_t, c_int, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_int]
        lib.or_functional.restype = c_double
        lib.or_functional.argtypes = [c_int, c_int, c_int, POINTER(c_double), c_size_t,
                                      c_void_p, c_size_t, POINTER(c_double), POINTER(c_double),
                                      c_size_t, c_double, c_void_p, c_int]
        lib.or_pull_back.restype = c_int
        lib.or_pull_back.argtypes = [c_int, c_int, c_int, c_int, POINTER(c_double), c_size_t,
                                     c_void_p, c_size_t, POINTER(c_double), c_size_t, c_int,
                                     c_double, c_void_p, c_void_p]
        lib.or_rel_err.restype = c_double
        lib.or_rel_err.argtypes = [POINTER(c_double), POINTER(c_double), c_size_t]
        _lib = lib
    return _lib


def _d(a):
    return a.ctypes.data_as(POINTER(c_double))


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


TDIM = {1: 1, 2: 2, 3: 2, 4: 3, 5: 3, 6: 3, 7: 3}


def sub_entities(cell: int, dim: int):
    """[(type, [vertices])] of the reference cell's sub-entities of dimension dim."""
    lib = load()
    out = []
    ty = c_int()
    verts = (c_int * 8)()
    while True:
        n = lib.or_sub_entity(cell, dim, len(out), ctypes.byref(ty), verts)
        if n < 0:
            return out
        out.append((ty.value, list(verts[:n])))


def lagrange_points(cell: int, degree: int, family: int = 0) -> np.ndarray:
    lib = load()
    k = lib.or_num_nodes(cell, degree)
    out = np.zeros(k * 3)
    n = lib.or_lagrange_points(cell, degree, family, _d(out))
    if n < 0:
        raise ValueError("unsupported element")
    td = TDIM[cell]
    return out[: n * td].reshape(n, td)


def tabulate(cell: int, degree: int, family: int, pts: np.ndarray):
    lib = load()
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    td = TDIM[cell]
    nq = pts.shape[0]
    k = lib.or_num_nodes(cell, degree)
    v = np.zeros(nq * k)
    d = np.zeros(nq * k * td)
    if lib.or_tabulate(cell, degree, family, _d(pts), nq, _p(v), _p(d)):
        raise ValueError("tabulation failed")
    return v.reshape(nq, k), d.reshape(nq, k, td)


def eval_block(cell, degree, family, gdim, nodes, dofmap, pts, factors=("points", "jacobians",
               "detj"), flags=None, ref_normals=None, threads=0):
    """Cell-major [c][q][LEN] f64 factors of one block at absolute f64 node coordinates."""
    lib = load()
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    td = TDIM[cell]
    k = lib.or_num_nodes(cell, degree)
    nc = dofmap.size // k
    nq = pts.shape[0]
    lens = {"points": gdim, "jacobians": td * gdim, "detj": 1, "pinv": td * gdim,
            "facet_normals": gdim, "normals": gdim}
    out = {f: np.zeros((nc, nq, lens[f])) for f in factors}
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
    rn = None if ref_normals is None else np.ascontiguousarray(ref_normals, dtype=np.float64)
    rc = lib.or_eval_block(cell, degree, family, gdim, _d(nodes), nodes.shape[0], _p(dofmap), nc,
                           _p(fl), _d(pts), nq, _p(rn), _p(out.get("points")),
                           _p(out.get("jacobians")), _p(out.get("detj")), _p(out.get("pinv")),
                           _p(out.get("facet_normals")), _p(out.get("normals")), threads)
    if rc:
        raise ValueError("oracle evaluation failed")
    return out


def cpu_factors(tdim, gdim, k, nodes_rel, origin, dofmap, values, derivs, tile=4, threads=0):
    """The timed CPU path: E-vector then X, J, det J into tiles (f64)."""
    lib = load()
    nodes_rel = np.ascontiguousarray(nodes_rel, dtype=np.float64)
    origin = np.ascontiguousarray(origin, dtype=np.float64)
    dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32)
    values = np.ascontiguousarray(values, dtype=np.float64)
    derivs = np.ascontiguousarray(derivs, dtype=np.float64)
    nc = dofmap.size // k
    nq = values.shape[0]
    padded = (nc + tile - 1) // tile * tile
    ev = np.empty(k * gdim * padded)
    x = np.empty(padded * nq * gdim)
    j = np.empty(padded * nq * tdim * gdim)
    dj = np.empty(padded * nq)
    lib.or_cpu_factors(tdim, gdim, k, _d(nodes_rel), nodes_rel.shape[0], _d(origin), _p(dofmap),
                       nc, _d(values), _d(derivs), nq, tile, _p(ev), _p(x), _p(j), _p(dj),
                       threads)
    return x, j, dj


def functional(cell, degree, family, nodes, dofmap, pts, wts, beta=0.1, grad=True):
    lib = load()
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32)
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    wts = np.ascontiguousarray(wts, dtype=np.float64)
    k = lib.or_num_nodes(cell, degree)
    g = np.zeros_like(nodes) if grad else None
    f = lib.or_functional(cell, degree, family, _d(nodes), nodes.shape[0], _p(dofmap),
                          dofmap.size // k, _d(pts), _d(wts), len(wts), beta, _p(g), 0)
    return f, g


def pull_back(cell, degree, family, gdim, nodes, dofmap, x, max_it=16, tol=1e-13):
    """Reference points xi [c][q][tdim] and last steps [c][q] of the physical points x
    [c][q][gdim] (absolute), by the Newton iteration of or_pull_back (pull_back.rs:26-180)."""
    lib = load()
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    td = lib.or_tdim(cell)
    c, nq = x.shape[0], x.shape[1]
    xi = np.zeros((c, nq, td))
    step = np.zeros((c, nq))
    rc = lib.or_pull_back(cell, degree, family, gdim, _d(nodes), nodes.shape[0], _p(dofmap), c,
                          _d(x), nq, max_it, tol, _p(xi), _p(step))
    if rc:
        raise ValueError("or_pull_back failed")
    return xi, step


_port = None


def cpu_port_lib() -> ctypes.CDLL:
    """The CPU-baseline port (cpu_port.cpp): the AVX-512 build where the host CPU has AVX-512F,
    else the AVX2/FMA build."""
    global _port
    if _port is None:
        flags = ""
        try:
            with open("/proc/cpuinfo") as f:
                flags = next((ln for ln in f if ln.startswith("flags")), "")
        except OSError:
            pass
        name = "libcpuport_v4.so" if " avx512f" in flags else "libcpuport_v3.so"
        path = os.path.join(HERE, "_build", name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        lib.cpu_port_factors.restype = c_int
        lib.cpu_port_factors.argtypes = [c_int, ctypes.c_uint, POINTER(c_double), c_size_t,
                                         POINTER(c_double),
                                         c_void_p, c_size_t, POINTER(c_double), POINTER(c_double),
                                         c_size_t, c_void_p, c_void_p, c_void_p, c_void_p, c_int]
        lib.isa = name
        _port = lib
    return _port


class CpuPort:
    """The port's buffers for one tetrahedral block (k = 4, 10 or 20) at one point set; call()
    runs the passes named by `passes` (bit 0 E-vector, 1 X, 2 J, 3 det J; outputs in tiles of 4:
    [tile][q][len][4])."""

    EVECTOR, X, J, DETJ, ALL = 1, 2, 4, 8, 15

    def __init__(self, k, nodes_rel, origin, dofmap, values, derivs, threads=0):
        self.lib = cpu_port_lib()
        self.k = k
        self.nodes = np.ascontiguousarray(nodes_rel, dtype=np.float64)
        self.origin = np.ascontiguousarray(origin, dtype=np.float64)
        self.dofmap = np.ascontiguousarray(dofmap, dtype=np.uint32).reshape(-1)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.derivs = np.ascontiguousarray(derivs, dtype=np.float64)
        self.cells = self.dofmap.size // k
        self.nq = self.values.shape[0]
        padded = (self.cells + 3) // 4 * 4
        self.ev = np.empty(3 * k * padded)
        self.x = np.empty(padded * self.nq * 3)
        self.j = np.empty(padded * self.nq * 9)
        self.detj = np.empty(padded * self.nq)
        self.threads = threads

    def __call__(self, passes=15):
        rc = self.lib.cpu_port_factors(self.k, passes, _d(self.nodes), self.nodes.shape[0],
                                       _d(self.origin), _p(self.dofmap), self.cells,
                                       _d(self.values), _d(self.derivs), self.nq, _p(self.ev),
                                       _p(self.x), _p(self.j), _p(self.detj), self.threads)
        if rc:
            raise ValueError("cpu_port_factors: unsupported element")
        return self.x, self.j, self.detj


def rel_err(got, want) -> float:
    got = np.ascontiguousarray(got, dtype=np.float64).reshape(-1)
    want = np.ascontiguousarray(want, dtype=np.float64).reshape(-1)
    assert got.size == want.size
    return load().or_rel_err(_d(got), _d(want), got.size)
