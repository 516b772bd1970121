"""Build recipe of the in-tree HIP library (gfx950): ``python -m gridcl_amd.build``.

Every HIP translation unit is compiled with packed FP32 instructions disabled; the host-only
element library is plain C++17. The result is ``gridcl_amd/lib/libtetraloma_geom.so``.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
ARCH = "gfx950"

# (source, extra defines): the factor kernels compile once per coordinate type.
HIP_SOURCES = [
    ("factor_kernels.hip", ["-DTG_FACTOR_TYPE=float", "-DTG_FACTOR_COMMON"]),
    ("factor_kernels.hip", ["-DTG_FACTOR_TYPE=double"]),
    # FP32 arithmetic on FP64-stored nodes
    ("factor_kernels.hip", ["-DTG_FACTOR_TYPE=float", "-DTG_FACTOR_NODES64=1"]),
    ("sumfact_kernels.hip", []),
    # The matrix-core kernel's products go straight to VGPRs (not AGPRs): its J is read by the
    # vector ALU at once, and the accumulation registers would otherwise count twice (4 waves
    # per SIMD instead of 3 for Q3).
    ("sumfact_mfma.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("dense_mfma.hip", ["-mllvm", "-amdgpu-mfma-vgpr-form"]),
    ("adjoint_kernels.hip", []),
    ("measure_kernels.hip", []),
    ("measure_noalloc.hip", []),
    ("pullback_kernels.hip", []),
    ("equispaced_kernels.hip", []),
]
HOST_HIP_SOURCES = ["capi.cpp", "adjoint.cpp", "plan.cpp", "tuning.cpp", "measure.cpp", "pullback.cpp",
                    "equispaced.cpp", "vtk_geometry.cpp"]  # host code that calls the HIP runtime
HOST_SOURCES = ["refcell.cpp", "gmsh_order.cpp", "gmsh_parse.cpp", "gmsh_mesh.cpp", "vtk_order.cpp",
                "vtk_write.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature; that warning is harmless.
    lines = [ln for ln in res.stdout.splitlines() if "packed-fp32-ops" not in ln]
    if res.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build_library(verbose: bool = False) -> str:
    """Compiles every source (in parallel) and links the shared library; returns its path."""
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = []
    for src, defs in HIP_SOURCES:
        tag = "".join(d.split("=")[-1][:1] for d in defs if "=" in d)
        obj = os.path.join(OBJ_DIR, f"{src}{'.' + tag if tag else ''}.o")
        jobs.append((obj, [HIPCC, *HIP_FLAGS, *defs, "-c", os.path.join(CSRC, src), "-o", obj]))
    for src in HOST_HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, src + ".o")
        jobs.append((obj, [HIPCC, "-O3", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, src),
                           "-o", obj]))
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ_DIR, src + ".o")
        jobs.append((obj, ["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-c",
                           os.path.join(CSRC, src), "-o", obj]))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd) for _, cmd in jobs]:
            f.result()
    objs = [obj for obj, _ in jobs]
    out = os.path.join(OUT_DIR, "libtetraloma_geom.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-lz", "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build_library(verbose=True)
