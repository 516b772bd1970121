"""The FP64 kernels against the committed per-configuration vectors of the FP64 oracle
(tests/golden/config_vectors.npz; needs the GPU). The sampled cells of each benchmark mesh form a
small geometry of their own; config 3's hexahedra at their 4^3 tensor points run the
sum-factorised kernel. Tolerance 1e-12 (rel_err, the FP64 bound of the other parity tests)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from gridcl_amd import api
from gridcl_amd import meshes as M

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "config_vectors.npz")


@pytest.mark.parametrize("c", [1, 2, 3, 5])
def test_fp64_kernels_match_golden_vectors(cuda, c):
    import torch

    d = np.load(GOLDEN)
    p = f"c{c}_"
    nodes = d[p + "nodes"]
    n, k, _ = nodes.shape
    cell, degree = int(d[p + "cell"]), int(d[p + "degree"])
    mesh = M.HostMesh(3, nodes.reshape(-1, 3).copy(),
                      [M.HostBlock(cell, degree, np.arange(n * k, dtype=np.uint32).reshape(n, k))])
    g = api.geometry_from_host(mesh, dtype=torch.float64)
    pts = d[p + "points"]
    ps = api.PointSet.custom(cell, pts, np.full(len(pts), 1.0 / len(pts)))
    fields = api.geometric_factors_block(g, 0, ps, ["points", "jacobians", "detj"])
    torch.cuda.synchronize()
    for name, key in (("points", "x"), ("jacobians", "j"), ("detj", "detj")):
        got = fields[name].to_host_cell_major()
        want = d[p + key].reshape(got.shape)
        err = O.rel_err(got, want)
        print(f"config {c} {name}: rel_err {err:.2e}")
        assert err <= 1e-12, (c, name, err)
