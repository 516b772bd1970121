"""Host element data against the reference's fixtures and against closed forms (CPU only).

Mirrors tetraloma-refcell/tests/fixtures.rs:52-193 (cell tables and Lagrange node layouts from
ndelement), tests/tabulation.rs:94-151 (Kronecker, partition of unity, reproduction) and
src/quadrature.rs:162-277 (rule volumes, exactness, interior points, ordering)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from gridcl_amd import meshes as M
from gridcl_amd.api import CellType, CoordinateElement, PointSet

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CELLS = {"Interval": 1, "Triangle": 2, "Quadrilateral": 3, "Tetrahedron": 4, "Hexahedron": 5,
         "Prism": 6, "Pyramid": 7}
FAMILIES = {"Equispaced": 0, "Gll": 1}


def _fixture(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_lagrange_points_match_fixture():
    """Library and oracle node layouts equal the ndelement export to 1e-14 (fixtures.rs:131)."""
    fx = _fixture("lagrange_points.json")
    count = 0
    for cname, fams in fx.items():
        for fname, degs in fams.items():
            for deg, vals in degs.items():
                want = np.array(vals)
                c, f, p = CELLS[cname], FAMILIES[fname], int(deg)
                lib = M.reference_nodes(c, p, f).reshape(-1)
                ora = O.lagrange_points(c, p, f).reshape(-1)
                assert lib.shape == want.shape and ora.shape == want.shape
                assert np.abs(lib - want).max() <= 1e-14, (cname, fname, deg)
                assert np.abs(ora - want).max() <= 1e-14, (cname, fname, deg)
                count += 1
    assert count >= 30


def test_vertices_match_tables_fixture():
    """Degree-1 nodes are the reference vertices of tables.json (cell.rs:53)."""
    tables = _fixture("tables.json")
    for cname, t in tables.items():
        if cname == "Point":
            continue
        want = np.array(t["vertices"], dtype=float).reshape(-1)
        c = CELLS[cname]
        assert np.array_equal(M.reference_nodes(c, 1).reshape(-1), want), cname
        assert np.array_equal(O.lagrange_points(c, 1).reshape(-1), want), cname


def test_sub_entities_match_tables_fixture():
    """The oracle's sub-entity tables (types and vertex lists, per dimension, in order) equal the
    reference's tables.json export (cell.rs:53-249). The A.5 node order walks these tables, so
    a reordered edge or face of any cell fails here."""
    tables = _fixture("tables.json")
    names = {0: "Point", 1: "Interval", 2: "Triangle", 3: "Quadrilateral", 4: "Tetrahedron",
             5: "Hexahedron", 6: "Prism", 7: "Pyramid"}
    ids = {v: k for k, v in names.items()}
    for cname, t in tables.items():
        c = ids[cname]
        assert O.load().or_tdim(c) == t["tdim"]
        for dim in range(t["tdim"] + 1):
            got = O.sub_entities(c, dim)
            want_v = t["sub_entities"][dim]
            want_t = t["sub_entity_types"][dim]
            assert [v for _, v in got] == want_v, (cname, dim)
            assert [names[ty] for ty, _ in got] == want_t, (cname, dim)


def test_node_counts():
    """A.4 counts and the doc examples of lagrange.rs:45-50."""
    for c in range(1, 8):
        for p in range(1, 6):
            assert len(M.reference_nodes(c, p)) == O.load().or_num_nodes(c, p)
    assert len(M.reference_nodes(6, 3)) == 40
    assert len(M.reference_nodes(7, 3)) == 30


def test_prism_pyramid_nodes_on_entities():
    """Prism and pyramid nodes (no fixture) lie in the reference cell, are distinct, and the
    ones of the degree-p lattice are multiples of 1/p (lagrange.rs:392-439)."""
    for c, pmax in ((6, 5), (7, 4)):
        for p in range(1, pmax + 1):
            x = M.reference_nodes(c, p)
            assert np.allclose(x * p, np.rint(x * p), atol=1e-12)
            assert len({tuple(np.rint(r * p).astype(int)) for r in x}) == len(x)
            if c == 6:
                assert np.all(x[:, 0] + x[:, 1] <= 1 + 1e-12)
            else:
                assert np.all(np.maximum(x[:, 0], x[:, 1]) + x[:, 2] <= 1 + 1e-12)
            assert np.array_equal(x, O.lagrange_points(c, p))


ELEMENTS = ([(1, p, f) for p in range(1, 9) for f in (0, 1)]
            + [(2, p, 0) for p in range(1, 9)]
            + [(3, p, f) for p in range(1, 9) for f in (0, 1)]
            + [(4, p, 0) for p in range(1, 9)]
            + [(5, p, f) for p in range(1, 6) for f in (0, 1)]
            + [(6, p, 0) for p in range(1, 6)]
            + [(7, p, 0) for p in range(1, 5)])


def _random_points(cell, n, seed=0):
    u = M.splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    td = M.TDIM[cell]
    if cell in (1, 3, 5):
        return u[:, :td]
    if cell == 2:
        a = u[:, :2]
        flip = a.sum(1) > 1
        a[flip] = 1 - a[flip]
        return a
    if cell == 4:
        s = np.sort(u, axis=1)
        return np.stack([s[:, 0], s[:, 1] - s[:, 0], s[:, 2] - s[:, 1]], axis=1)
    if cell == 6:
        a = u[:, :2].copy()
        flip = a.sum(1) > 1
        a[flip] = 1 - a[flip]
        return np.concatenate([a, u[:, 2:3]], axis=1)
    z = 0.95 * u[:, 2]
    return np.stack([u[:, 0] * (1 - z), u[:, 1] * (1 - z), z], axis=1)


def _monomials(cell, p):
    """Exponents of polynomials inside the element's space."""
    out = []
    for a in range(p + 1):
        for b in range(p + 1):
            for c in range(p + 1):
                if cell in (2, 4) and a + b + c > p:
                    continue
                if cell == 6 and (a + b > p or c > p):
                    continue
                if cell == 7 and (a + b + c > p):
                    continue
                td = M.TDIM[cell]
                if (td < 2 and b) or (td < 3 and c):
                    continue
                out.append((a, b, c))
    return out


@pytest.mark.parametrize("cell,p,f", ELEMENTS)
def test_oracle_basis_properties(cell, p, f):
    """Kronecker at the nodes, partition of unity, and reproduction of the element's
    polynomials and their gradients (tests/tabulation.rs:94-151)."""
    nodes = O.lagrange_points(cell, p, f)
    v, _ = O.tabulate(cell, p, f, nodes)
    assert np.abs(v - np.eye(len(nodes))).max() < 1e-12
    x = _random_points(cell, 17, seed=cell * 10 + p)
    v, d = O.tabulate(cell, p, f, x)
    assert np.abs(v.sum(1) - 1).max() < 1e-11
    assert np.abs(d.sum(1)).max() < 1e-9
    td = M.TDIM[cell]
    for a, b, c in _monomials(cell, p)[::3]:
        e = np.array([a, b, c][:td])
        fn = np.prod(nodes ** e, axis=1)
        want = np.prod(x ** e, axis=1)
        assert np.abs(v @ fn - want).max() < 1e-10, (a, b, c)
        for dd in range(td):
            ed = e.copy()
            coef = ed[dd]
            ed[dd] = max(ed[dd] - 1, 0)
            gw = coef * np.prod(x ** ed, axis=1)
            assert np.abs(d[:, :, dd] @ fn - gw).max() < 1e-8, (a, b, c, dd)


@pytest.mark.parametrize("cell,p,f", ELEMENTS)
def test_library_tables_match_oracle(cell, p, f):
    """The library's Vandermonde tabulation equals the oracle's closed forms."""
    ps = PointSet.gauss(cell, 2 * p)
    e = CoordinateElement(CellType(cell), p, f)
    v, d = ps.tabulate(e), ps.tabulate_derivatives(e)
    ov, od = O.tabulate(cell, p, f, ps.points)
    # both sides solve a Vandermonde system for pyramids; at p = 4 its conditioning costs a digit
    tol = 1e-11 if (cell, p) == (7, 4) else 1e-12
    assert np.abs(v - ov).max() < tol
    assert np.abs(d - od).max() < 10 * tol


def _exact(cell, a, b, c):
    """Exact integral of x^a y^b z^c over the reference cell."""
    f = math.factorial
    if cell == 1:
        return 1.0 / (a + 1)
    if cell == 3:
        return 1.0 / ((a + 1) * (b + 1))
    if cell == 5:
        return 1.0 / ((a + 1) * (b + 1) * (c + 1))
    if cell == 2:
        return f(a) * f(b) / f(a + b + 2)
    if cell == 4:
        return f(a) * f(b) * f(c) / f(a + b + c + 3)
    if cell == 6:
        return f(a) * f(b) / f(a + b + 2) / (c + 1)
    raise ValueError


VOLUME = {1: 1.0, 2: 0.5, 3: 1.0, 4: 1.0 / 6, 5: 1.0, 6: 0.5, 7: 1.0 / 3}


@pytest.mark.parametrize("cell", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("degree", [0, 1, 2, 3, 4, 5, 7])
def test_gauss_rules(cell, degree):
    """m^tdim points (m = degree/2 + 1), positive weights summing to the volume, interior
    points, first coordinate fastest, exact on the cell's monomials (quadrature.rs:162-277)."""
    ps = PointSet.gauss(cell, degree)
    m = degree // 2 + 1
    td = M.TDIM[cell]
    assert len(ps) == m ** td
    assert np.all(ps.weights > 0)
    assert abs(ps.weights.sum() - VOLUME[cell]) < 1e-14
    x = ps.points
    assert np.all(x > 0) and np.all(x < 1)
    if cell in (3, 5) and m > 1:
        assert x[0, 0] < x[1, 0] and abs(x[0, 1] - x[1, 1]) < 1e-15
    if cell == 7:
        return
    for a in range(degree + 1):
        for b in range(degree + 1 if td > 1 else 1):
            for c in range(degree + 1 if td > 2 else 1):
                if cell in (2, 4) and a + b + c > degree:
                    continue
                if cell == 6 and a + b > degree:
                    continue
                e = np.array([a, b, c][:td])
                got = ps.weights @ np.prod(x ** e, axis=1)
                assert abs(got - _exact(cell, a, b, c)) < 1e-14, (a, b, c)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_pyramid_basis_reproduces_rational_space(p):
    """The oracle's pyramid basis of degree p reproduces every member x^a y^b z^c / (1-z)^min(a,b)
    (a + c <= p, b + c <= p) of the reference's pyramid space (polynomials/pyramid.rs:1-9) and
    its gradient, including the rational ones (min(a, b) >= 1), which no polynomial check
    reaches. The members are counted against A.4's dimension (p+1)(p+2)(2p+3)/6."""
    nodes = O.lagrange_points(7, p)
    members = [(a, b, c) for c in range(p + 1) for a in range(p + 1 - c)
               for b in range(p + 1 - c)]
    assert len(members) == len(nodes) == (p + 1) * (p + 2) * (2 * p + 3) // 6
    x = _random_points(7, 23, seed=70 + p)
    v, d = O.tabulate(7, p, 0, x)
    for a, b, c in members:
        m = min(a, b)

        def f(y):
            s = 1 - y[:, 2]
            out = np.zeros(len(y))
            ok = s > 0
            out[ok] = y[ok, 0] ** a * y[ok, 1] ** b * y[ok, 2] ** c / s[ok] ** m
            # apex: x = y = 0, the limit is 1 for a = b = 0 and 0 otherwise
            out[~ok] = 1.0 if a == 0 and b == 0 else 0.0
            return out

        fn = f(nodes)
        s = 1 - x[:, 2]
        xa, yb, zc = x[:, 0] ** a, x[:, 1] ** b, x[:, 2] ** c
        gx = a * x[:, 0] ** max(a - 1, 0) * yb * zc / s ** m
        gy = b * xa * x[:, 1] ** max(b - 1, 0) * zc / s ** m
        gz = xa * yb * (c * x[:, 2] ** max(c - 1, 0) / s ** m + m * zc / s ** (m + 1))
        assert np.abs(v @ fn - f(x)).max() < 1e-10, (a, b, c)
        for dd, g in enumerate((gx, gy, gz)):
            assert np.abs(d[:, :, dd] @ fn - g).max() < 1e-8, (a, b, c, dd)


def test_pyramid_rule_exact_on_rational_space():
    """The pyramid rule of degree 2p integrates products of two degree-p basis functions' space
    members exactly: check the mass of x^a y^b z^c/(1-z)^min(a,b) against a fine rule."""
    coarse = PointSet.gauss(7, 4)
    fine = PointSet.gauss(7, 30)
    for a, b, c in [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 0, 2), (2, 1, 0), (1, 1, 1), (2, 2, 0)]:
        def f(x):
            s = 1 - x[:, 2]
            return x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c / s ** min(a, b)
        assert abs(coarse.weights @ f(coarse.points) - fine.weights @ f(fine.points)) < 1e-13


def test_tetrahedron_rules_not_in_reference():
    """Keast's 4-point rule is exact to degree 2 and the 14-point rule to degree 5, with
    positive weights summing to 1/6 and interior points; neither is exact one degree higher."""
    for ps, deg in ((PointSet.tet_keast4(), 2), (PointSet.tet_14(), 5)):
        assert np.all(ps.weights > 0) and abs(ps.weights.sum() - 1 / 6) < 1e-15
        x = ps.points
        assert np.all(x > 0) and np.all(x.sum(1) < 1)
        worst_above = 0.0
        for a in range(deg + 2):
            for b in range(deg + 2 - a):
                for c in range(deg + 2 - a - b):
                    got = ps.weights @ (x[:, 0] ** a * x[:, 1] ** b * x[:, 2] ** c)
                    err = abs(got - _exact(4, a, b, c))
                    if a + b + c <= deg:
                        assert err < 1e-15, (a, b, c)
                    else:
                        worst_above = max(worst_above, err)
        assert worst_above > 1e-6
    assert len(PointSet.tet_keast4()) == 4 and len(PointSet.tet_14()) == 14


def test_facet_rules_and_reference_normals():
    """Facet point sets lie on their facets with the outward normal of each facet
    (facet.rs:46-92; facets.rs:27-42: the unit square's four axis normals)."""
    from gridcl_amd import api

    fq = api.PointSet.facet_gauss(3, 1)
    assert len(fq) == 4
    fr = api.PointSet.facet_gauss(2, 3)
    assert len(fr) == 6  # two Gauss points per triangle edge (facet_rule.rs:30-35)
    t = api.PointSet.facet_gauss(4, 2)
    assert len(t) == 4 * 4
    x = t.points
    # Facet 3 of the tetrahedron is (1,2,3): x + y + z = 1.
    assert np.allclose(x[12:].sum(1), 1.0)
    assert np.allclose(x[:4, 2], 0.0)
