"""Gmsh MSH 4.1 reading (tetraloma-io/src/gmsh; tetraloma-refcell/src/node_maps/gmsh.rs) on the
host: the reference's fixture corpus (tetraloma-io/tests/fixtures/gmsh, written by Gmsh 4.15.2,
copied under tests/golden/gmsh with its manifest of Gmsh's own counts).

Pinned by: the manifest (node and element counts per type, dimensions, physical groups with the
sorted corner tags of every element, duplicate ids), the geometry of the corpus (every cell is an
affine image of its reference cell, so each node read in A.5 order must sit at the affine image
of its A.5 reference point: a wrong Gmsh-to-A.5 permutation fails), the complete-P2 twins of the
serendipity files, the binary copies of the ASCII files, the files re-encoded in both byte orders
and both size_t widths, the reference's doc-test examples and its error rules (as in
tetraloma-io/tests/gmsh_corpus.rs and tetraloma/tests/gmsh.rs)."""
import json
import os

import numpy as np
import pytest

from gridcl_amd import api, gmsh
from gridcl_amd import meshes as M

from tests import msh_sections

DIR = os.path.join(os.path.dirname(__file__), "golden", "gmsh")
MANIFEST = json.load(open(os.path.join(DIR, "manifest.json")))
FILES = MANIFEST["files"]
MESHES = [f for f in FILES if "expect_error" not in f]
NV = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5}


def path(name):
    return os.path.join(DIR, name)


def corners(mesh, b):
    """Corner node tags of the cells of block b, [cell][vertices]."""
    blk = mesh.blocks[b]
    return mesh.node_tags[blk.nodes[:, :NV[blk.cell]]]


def test_doc_examples():
    """convert.rs:21-33, 99-113, mod.rs:28-39: two triangles and a tag edge."""
    text = (b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
            b"$Nodes\n1 4 1 4\n2 1 0 4\n1\n2\n3\n4\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n$EndNodes\n"
            b"$Elements\n2 3 1 3\n2 1 2 2\n1 1 2 3\n2 1 3 4\n1 5 1 1\n3 1 3\n$EndElements\n")
    m = gmsh.parse(text)
    assert (m.tdim, m.gdim) == (2, 2)
    assert m.blocks[0].cell == M.TRIANGLE and m.blocks[0].degree == 1
    assert m.blocks[0].nodes.reshape(-1).tolist() == [0, 1, 2, 0, 2, 3]
    assert m.blocks[0].tags.tolist() == [1, 2] and m.blocks[0].elementary.tolist() == [1, 1]
    assert m.node_tags.tolist() == [1, 2, 3, 4] and m.nodes[2].tolist() == [1.0, 1.0]
    g = m.tag_groups[0]
    assert (g.dim, g.cell) == (1, M.INTERVAL)
    assert g.corners.reshape(-1).tolist() == [1, 3] and g.tags.tolist() == [3]
    assert g.elementary.tolist() == [5]
    # node_maps.rs:31-42: Gmsh's quadrilateral runs round the boundary.
    assert gmsh.gmsh_to_internal(M.QUADRILATERAL, 1).tolist() == [0, 1, 3, 2]
    # parse.rs:552-565 and model.rs:11-17: names, entities.
    m2 = gmsh.parse(b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
                    b"$PhysicalNames\n1\n2 7 \"domain\"\n$EndPhysicalNames\n"
                    b"$Nodes\n1 3 1 3\n2 1 0 3\n1\n2\n3\n0 0 0\n1 0 0\n0 1 0\n$EndNodes\n"
                    b"$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 3\n$EndElements\n")
    assert m2.names == [(2, 7, "domain")]
    assert m2.element_blocks == [(2, 1, 2, 1)]


@pytest.mark.parametrize("f", MESHES, ids=[f["file"] for f in MESHES])
def test_corpus_counts_and_dimensions(f):
    """Gmsh's node count, element counts per type, tdim, gdim, encoding, partitioned and periodic
    flags (gmsh_corpus.rs:29-61; tetraloma/tests/gmsh.rs:20-57, 115-172), and the distinct nodes
    of the cells (for serendipity files: those of the complete-P2 twin)."""
    m = gmsh.read(path(f["file"]))
    assert m.file_nodes == f["num_nodes"]
    assert m.element_counts == {int(k): v for k, v in f["element_counts"].items()}
    assert (m.tdim, m.gdim) == (f["tdim"], f["gdim"])
    assert m.binary == f["binary"]
    assert m.partitioned == f["partitioned"]
    assert m.num_partitions == f["num_partitions"]
    assert (m.num_periodic > 0) == f["periodic"]
    used = np.unique(np.concatenate([b.nodes.reshape(-1) for b in m.blocks]))
    assert len(used) == f["num_mesh_nodes"]
    assert m.nodes.shape == (len(m.node_tags), m.gdim)


@pytest.mark.parametrize("f", MESHES, ids=[f["file"] for f in MESHES])
def test_physical_groups(f):
    """Each physical group's elements, as sorted corner tags, and its name are Gmsh's
    (tetraloma/tests/gmsh.rs:59-112)."""
    m = gmsh.read(path(f["file"]))
    names = {(d, t): n for d, t, n in m.names}
    for grp in f["physical_groups"]:
        dim, tag = grp["dim"], grp["tag"]
        got = []
        if dim == m.tdim:
            for b, blk in enumerate(m.blocks):
                cells = blk.physical[blk.physical[:, 1] == tag, 0]
                cs = corners(m, b)
                got += [sorted(cs[c].tolist()) for c in cells]
        else:
            for g in m.tag_groups:
                if g.dim != dim:
                    continue
                idx = g.physical[g.physical[:, 1] == tag, 0]
                got += [sorted(g.corners[e].tolist()) for e in idx]
        assert len(got) == grp["size"], (dim, tag)
        assert sorted(got) == sorted(grp["elements"]), (dim, tag)
        assert names.get((dim, tag)) == grp["name"]


def affine_fit(ref_vertices, x_vertices):
    """The affine map ref -> x through the vertices (least squares) and its worst residual."""
    a = np.concatenate([ref_vertices, np.ones((len(ref_vertices), 1))], axis=1)
    coef, *_ = np.linalg.lstsq(a, x_vertices, rcond=None)
    return coef, np.abs(a @ coef - x_vertices).max()


AFFINE = [f for f in MESHES if f["file"].split("_")[0] in
          ("interval", "triangle", "quadrilateral", "tetrahedron", "hexahedron", "prism", "hybrid")]


@pytest.mark.parametrize("f", AFFINE, ids=[f["file"] for f in AFFINE])
def test_nodes_sit_at_affine_images_of_a5_points(f):
    """Every corpus cell is an affine image of its reference cell, so node k (A.5 order) of a
    degree-p cell lies at F(xi_k) with F through the cell's vertices and xi_k the k-th A.5
    equispaced point: pins the Gmsh-to-A.5 permutation of every cell type and degree in the
    corpus (degrees 1-4, pyramids 1-3; serendipity files after completion)."""
    m = gmsh.read(path(f["file"]))
    scale = np.abs(m.nodes).max()
    for blk in m.blocks:
        ref = M.reference_nodes(blk.cell, blk.degree)
        nv = NV[blk.cell]
        rv = ref[:nv]
        for c in range(blk.nodes.shape[0]):
            x = m.nodes[blk.nodes[c]]
            coef, res = affine_fit(rv, x[:nv])
            assert res <= 1e-12 * scale
            want = np.concatenate([ref, np.ones((len(ref), 1))], axis=1) @ coef
            assert np.abs(want - x).max() <= 1e-10 * scale, (blk.cell, blk.degree, c)


SERENDIPITY = [f for f in MESHES if f.get("twin")]


@pytest.mark.parametrize("f", SERENDIPITY, ids=[f["file"] for f in SERENDIPITY])
def test_serendipity_reads_as_its_p2_twin(f):
    """A serendipity file completes to the mesh of its complete-P2 twin: the same cells with the
    same coordinates at every A.5 node (the added nodes are the serendipity interpolant, exact on
    affine cells), and the same number of distinct nodes (face nodes shared)."""
    s = gmsh.read(path(f["file"]))
    t = gmsh.read(path(f["twin"]))
    assert [(b.cell, b.degree) for b in s.blocks] == [(b.cell, b.degree) for b in t.blocks]
    for bs, bt in zip(s.blocks, t.blocks):
        assert (bs.tags == bt.tags).all()
        assert np.abs(s.nodes[bs.nodes] - t.nodes[bt.nodes]).max() <= 1e-14
    used = lambda m: len(np.unique(np.concatenate([b.nodes.reshape(-1) for b in m.blocks])))
    assert used(s) == used(t)
    # Added nodes follow the file's, with tags above the largest file tag.
    assert len(s.node_tags) > s.file_nodes
    assert s.node_tags[s.file_nodes:].min() > s.node_tags[:s.file_nodes].max()


def same_mesh(a, b, tol=0.0):
    """Equal meshes; coordinates within `tol` (Gmsh's ASCII files hold 16 significant digits,
    its binary files the doubles themselves)."""
    assert (a.tdim, a.gdim) == (b.tdim, b.gdim)
    assert a.nodes.shape == b.nodes.shape and np.abs(a.nodes - b.nodes).max(initial=0) <= tol
    assert np.array_equal(a.node_tags, b.node_tags)
    assert len(a.blocks) == len(b.blocks) and len(a.tag_groups) == len(b.tag_groups)
    for x, y in zip(a.blocks, b.blocks):
        assert (x.cell, x.degree) == (y.cell, y.degree)
        for k in ("nodes", "tags", "elementary", "physical"):
            assert np.array_equal(getattr(x, k), getattr(y, k)), k
    for x, y in zip(a.tag_groups, b.tag_groups):
        assert (x.dim, x.cell) == (y.dim, y.cell)
        for k in ("corners", "tags", "elementary", "physical"):
            assert np.array_equal(getattr(x, k), getattr(y, k)), k
    assert a.names == b.names


BINARY = [f for f in MESHES if f["binary"]]


@pytest.mark.parametrize("f", BINARY, ids=[f["file"] for f in BINARY])
def test_binary_copies_read_as_their_ascii_files(f):
    """Gmsh's binary copy of a file gives the same mesh as the ASCII file (coordinates to the
    ASCII file's 16 digits)."""
    same_mesh(gmsh.read(path(f["file"])), gmsh.read(path(f["file"].replace("-binary", ""))),
              tol=4e-15)


ASCII = [f for f in MESHES if not f["binary"]]


@pytest.mark.parametrize("f", ASCII, ids=[f["file"] for f in ASCII])
def test_reencoded_files_read_the_same(f):
    """Every ASCII file, read section by section in plain Python and re-encoded as ASCII and as
    binary in both byte orders with 4- and 8-byte size_t, gives the same mesh
    (gmsh_corpus.rs:223-251); the plain reading agrees with the library's counts."""
    sec = msh_sections.read_ascii(path(f["file"]))
    want = gmsh.read(path(f["file"]))
    assert sum(len(b[2]) for b in sec.node_blocks) == want.file_nodes
    assert [(b[0], b[1], b[2], len(b[3])) for b in sec.element_blocks] == want.element_blocks
    for binary, big, size_t in [(False, False, 8), (True, False, 8), (True, False, 4),
                                (True, True, 8), (True, True, 4)]:
        got = gmsh.parse(gmsh.encode(sec, binary, big, size_t))
        same_mesh(got, want)
        assert got.binary == binary


@pytest.mark.parametrize("name,key", [("duplicate_node_tags.msh", "nodes"),
                                      ("duplicate_element_tags.msh", "elements")])
def test_duplicate_tags(name, key):
    f = next(x for x in FILES if x["file"] == name)
    assert f["duplicate_of"] == key
    with pytest.raises(api.DuplicateOriginId) as e:
        gmsh.read(path(name))
    assert e.value.detail == f["duplicate_id"]


def test_truncated_binary_files_fail_with_parse():
    """Every prefix of a binary file cut inside $Nodes .. $EndElements fails with Parse, in both
    byte orders, except the two prefixes that end right after $EndNodes (a file without
    elements): gmsh_corpus.rs:256-296."""
    sec = msh_sections.read_ascii(path("two_regions_p1.msh"))
    for big in (False, True):
        data = gmsh.encode(sec, True, big, 8)
        start = data.index(b"$Nodes\n") + 7
        end = data.index(b"$EndElements\n")
        between = data.index(b"$EndNodes\n") + 9
        for n in range(start, end):
            if n in (between, between + 1):
                continue
            with pytest.raises(api.Parse):
                gmsh.parse(data[:n])


def test_damaged_files_end_in_a_reader_error():
    """Seeded damage to corpus files (byte changes, cuts, insertions, truncations) never crashes
    the reader: each file parses or raises one of the reader's error variants, whose message is
    printable text (bytes of the file are quoted as escapes)."""
    import random

    rng = random.Random(2024)
    names = sorted(f["file"] for f in FILES)
    for _ in range(1500):
        data = bytearray(open(path(rng.choice(names)), "rb").read())
        for _ in range(rng.randint(1, 4)):
            if not data:
                break
            i = rng.randrange(len(data))
            op = rng.random()
            if op < 0.5:
                data[i] = rng.randrange(256)
            elif op < 0.7:
                del data[i:i + rng.randint(1, 16)]
            elif op < 0.85:
                data[i:i] = bytes(rng.randrange(256) for _ in range(rng.randint(1, 8)))
            else:
                del data[i:]
        try:
            gmsh.parse(bytes(data))
        except api.GridError as e:
            assert "\ufffd" not in str(e), str(e)  # valid text: the file's bytes came escaped


HEAD = b"$MeshFormat\n4.1 0 8\n$EndMeshFormat\n"
NODES = b"$Nodes\n1 3 1 3\n2 1 0 3\n1\n2\n3\n0 0 0\n1 0 0\n0 1 0\n$EndNodes\n"


def test_errors():
    """Error variants and their details (parse.rs:213-245, 409-500; convert.rs:183-206)."""
    with pytest.raises(api.Parse) as e:
        gmsh.parse(b"$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
    assert e.value.detail == 12
    with pytest.raises(api.Parse):
        gmsh.parse(b"hello")
    with pytest.raises(api.Parse):
        gmsh.parse(HEAD + b"$Nodes\n1 3 1 3\n2 1 0 3\n1\n2\n0\n0 0 0\n1 0 0\n0 1 0\n$EndNodes\n")
    with pytest.raises(api.Parse):  # announced total
        gmsh.parse(HEAD + b"$Nodes\n1 4 1 3\n2 1 0 3\n1\n2\n3\n0 0 0\n1 0 0\n0 1 0\n$EndNodes\n")
    with pytest.raises(api.UnsupportedElementType) as e:
        gmsh.parse(HEAD + NODES + b"$Elements\n1 1 1 1\n2 1 99 1\n1 1 2 3\n$EndElements\n")
    assert e.value.detail == 99
    with pytest.raises(api.Parse):  # a triangle in a block of dimension 3
        gmsh.parse(HEAD + NODES + b"$Elements\n1 1 1 1\n3 1 2 1\n1 1 2 3\n$EndElements\n")
    with pytest.raises(api.Parse):  # missing node
        gmsh.parse(HEAD + NODES + b"$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 4\n$EndElements\n")
    with pytest.raises(api.Parse):  # no elements of dimension >= 1
        gmsh.parse(HEAD + NODES + b"$Elements\n1 1 1 1\n0 1 15 1\n1 1\n$EndElements\n")
    with pytest.raises(api.DimensionMismatch):
        gmsh.parse(HEAD + NODES + b"$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 3\n$EndElements\n",
                   gdim=1)
    m = gmsh.parse(HEAD + NODES + b"$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 3\n$EndElements\n",
                   gdim=3)
    assert m.gdim == 3 and m.nodes.shape == (3, 3)
    six = (b"$Nodes\n1 6 1 6\n2 1 0 6\n1\n2\n3\n4\n5\n6\n0 0 0\n1 0 0\n0 1 0\n0.5 0 0\n"
           b"0.5 0.5 0\n0 0.5 0\n$EndNodes\n")
    with pytest.raises(api.DuplicateCellBlock):
        gmsh.parse(HEAD + six + b"$Elements\n2 2 1 2\n2 1 2 1\n1 1 2 3\n"
                   b"2 2 9 1\n2 1 2 3 4 5 6\n$EndElements\n")
    with pytest.raises(api.UnmatchedTagElement) as e:  # an edge through a mid-edge node
        gmsh.parse(HEAD + six + b"$Elements\n2 2 1 2\n2 1 2 1\n1 1 2 3\n"
                   b"1 1 1 1\n7 4 5\n$EndElements\n")
    assert e.value.detail == 7
    with pytest.raises(api.Io):
        gmsh.read(path("no-such-file.msh"))


def test_unknown_sections_and_comments_are_skipped():
    text = (HEAD + b"$Comments\nanything $Nodes here\n$EndComments\n" + NODES +
            b"$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 3\n$EndElements\n")
    m = gmsh.parse(text)
    assert m.blocks[0].nodes.tolist() == [[0, 1, 2]]


def test_partitioned_file_reads_as_its_serial_mesh():
    """Elements on partition interfaces are dropped and elementary tags are the parents'
    (chunk.rs:77-124): the partitioned file gives the cells of the unpartitioned one."""
    p = gmsh.read(path("partitioned_p1.msh"))
    s = gmsh.read(path("two_regions_p1.msh"))
    for bp, bs in zip(p.blocks, s.blocks):
        assert sorted(map(tuple, np.sort(corners(p, p.blocks.index(bp)), axis=1))) == \
            sorted(map(tuple, np.sort(corners(s, s.blocks.index(bs)), axis=1)))
        assert set(bp.elementary.tolist()) <= {e["tag"] for e in
                                               msh_sections.read_ascii(path("two_regions_p1.msh"))
                                               .entities[p.tdim]}


def test_node_maps_are_permutations():
    """gmsh_to_internal is a permutation for every cell up to its highest Gmsh degree, starts with
    the vertices, and matches the tables Gmsh's element definitions give at low degree."""
    top = {M.INTERVAL: 10, M.TRIANGLE: 10, M.QUADRILATERAL: 10, M.TETRAHEDRON: 10,
           M.HEXAHEDRON: 9, M.PRISM: 9, M.PYRAMID: 9}
    for cell, pmax in top.items():
        for p in range(1, pmax + 1):
            m = gmsh.gmsh_to_internal(cell, p)
            assert sorted(m.tolist()) == list(range(len(m)))
            assert sorted(m[:NV[cell]].tolist()) == list(range(NV[cell]))
    assert gmsh.gmsh_to_internal(M.HEXAHEDRON, 1).tolist() == [0, 1, 3, 2, 4, 5, 7, 6]
    assert gmsh.gmsh_to_internal(M.PYRAMID, 1).tolist() == [0, 1, 3, 2, 4]
    assert gmsh.gmsh_to_internal(M.TETRAHEDRON, 1).tolist() == [0, 1, 2, 3]
