"""tg_measure_block_noalloc, the measure whose work memory is a fixed array inside the library
(no allocation), on the GPU through the C ABI: bit for bit tg_measure_block's value on two calls
in a row, the second on a smaller block (the array keeps the first call's contents), on the
one-kernel schedule (register kernels; the last workgroup finishes the sums) and on the det J
field schedule (sum-factorised hexahedra); and its two refusals, made before anything runs."""
import numpy as np
import pytest

from gridcl_amd import _lib, api
from gridcl_amd import meshes as M

pytestmark = pytest.mark.gpu


def _call(fn, g, ps, options=0):
    import torch

    val = torch.full((), float("nan"), dtype=torch.float64, device="cuda")
    rc = fn(g.handle, 0, ps._h, options, val.data_ptr(), api._stream())
    torch.cuda.synchronize()
    return rc, float(val.item())


def _pair(lib, mesh, ps, tdt, options=0):
    g = api.geometry_from_host(mesh, dtype=tdt)
    rc, want = _call(lib.tg_measure_block, g, ps, options)  # also uploads the tables
    assert rc == 0, _lib.last_error()
    rc, got = _call(lib.tg_measure_block_noalloc, g, ps, options)
    assert rc == 0, _lib.last_error()
    return got, want


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_noalloc_equals_measure_twice(cuda, dtype):
    import torch

    lib = _lib.load()
    tdt = torch.float64 if dtype == "f64" else torch.float32
    ps = api.PointSet.tet_14()
    # 82,944 then 3,072 curved P2 tetrahedra: 1,296 then 48 first-pass partials.
    for n in (24, 8):
        got, want = _pair(lib, M.config2_mesh(n=n), ps, tdt)
        print(f"P2 tets n={n} {dtype}: noalloc {got!r} measure {want!r}")
        assert np.float64(got).tobytes() == np.float64(want).tobytes()
        # The curved box's volume, about 1 (the rule is not exact on these curved cells).
        assert abs(got - 1.0) <= 1e-2
    # Q2 hexahedra through the det J field schedule (sum factorisation forced): 6^3 then 3^3
    # cells x 27 points.
    hps = api.PointSet.gauss(M.HEXAHEDRON, 5)
    for n in (6, 3):
        mesh = M.jitter(M.cube_box(M.HEXAHEDRON, n, 2), 0.02 / n, 5, interior_only=True)
        got, want = _pair(lib, mesh, hps, tdt, api.OPT_SUMFACT_ALWAYS)
        print(f"Q2 hexes n={n} {dtype}: noalloc {got!r} measure {want!r}")
        assert np.float64(got).tobytes() == np.float64(want).tobytes()


def test_noalloc_refusals(cuda):
    """Tables not yet resident: TG_ERR_INVALID_ARGUMENT. A det J field beyond the array (Q3
    hexahedra 16^3 x 64 points in FP64: 2 MiB): TG_ERR_TOO_LARGE. The value stays untouched."""
    import torch

    lib = _lib.load()
    mesh = M.config2_mesh(n=4)
    g = api.geometry_from_host(mesh, dtype=torch.float32)
    fresh = api.PointSet.tet_14()
    rc, val = _call(lib.tg_measure_block_noalloc, g, fresh)
    assert rc == api.InvalidArgument.code and np.isnan(val), (rc, val)
    assert "not resident" in _lib.last_error()
    big = M.cube_box(M.HEXAHEDRON, 16, 3)
    gb = api.geometry_from_host(big, dtype=torch.float64)
    ps = api.PointSet.gauss(M.HEXAHEDRON, 7)
    rc, want = _call(lib.tg_measure_block, gb, ps)
    assert rc == 0 and abs(want - 1.0) <= 1e-12
    rc, val = _call(lib.tg_measure_block_noalloc, gb, ps)
    assert rc == api.TooLarge.code and np.isnan(val), (rc, val)
    assert "TooLarge" in _lib.last_error()
