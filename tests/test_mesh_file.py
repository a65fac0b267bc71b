"""MFEM v1.0 mesh files (Mesh::ReadMFEMMesh, mesh/mesh_readers.cpp:40-130):
the oracle's reader and the library's host-side loader.

Fixtures: tests/golden/meshes/{beam-hex,fichera,ref-cube}.mesh are copies of
MFEM's data/ files of the same names (mesh data, as the reference's own tests
read them).

Pinning: a Cartesian mesh written out as v1.0 text (in lexicographic or SFC
element order) must read back to exactly the topology of the pinned Cartesian
builder with the same element order; the Fichera corner's counts follow from
its geometry (26 vertices, 51 edges, 33 faces, 7 cells; 19 interior DOFs at
p = 2: the cells, 9 inner faces and 3 inner edges); every shared DOF of a
file mesh sits at one physical point.  The library's loader must give the
oracle's topology bit for bit.
"""
import numpy as np
import pytest

from test_oracle_mesh import _elem_lex_coords

MESHES = "tests/golden/meshes"


def _text(name):
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    with open(os.path.join(here, MESHES, name)) as f:
        return f.read()


def _cartesian_text(orc, dims, sfc):
    m = orc.CartesianMesh(*dims, sfc=sfc)
    xyz = m.vertices(0.0, 0)
    # boundary: the cube's faces in an arbitrary order (the file's order does
    # not matter), each with its own vertex cycle
    faces = {}
    hexf = [[3, 2, 1, 0], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7], [4, 5, 6, 7]]
    for e in range(m.ne):
        for j in range(6):
            f = m.ef[e, j]
            faces.setdefault(f, []).append([int(m.ev[e, k]) for k in hexf[j]])
    bdr = [v[0] for v in faces.values() if len(v) == 1][::-1]
    lines = ["MFEM mesh v1.0", "", "# written by the test", "dimension", "3", "",
             "elements", str(m.ne)]
    lines += ["1 5 " + " ".join(str(int(v)) for v in m.ev[e]) for e in range(m.ne)]
    lines += ["", "boundary", str(len(bdr))] + ["2 3 " + " ".join(map(str, q)) for q in bdr]
    lines += ["", "vertices", str(m.nv), "3"] + [" ".join(repr(float(c)) for c in x) for x in xyz]
    return "\n".join(lines) + "\n", m


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 2, 2), (3, 2, 4)])
@pytest.mark.parametrize("sfc", [False, True])
def test_cartesian_round_trip_matches_cartesian_builder(orc, dims, sfc):
    text, m = _cartesian_text(orc, dims, sfc)
    P = orc.Problem(None, 2, mesh_text=text)
    f = P.mesh
    assert (f.nv, f.ne, f.nedges, f.nfaces) == (m.nv, m.ne, m.nedges, m.nfaces)
    for k in ("ev", "ee", "eeo", "ef", "efo"):
        assert np.array_equal(getattr(f, k), getattr(m, k)), k
    ed, ndofs = m.elem_dofs(2)
    assert np.array_equal(P.ess, m.boundary_dofs(2, ed))


def test_fichera_counts(orc):
    P = orc.Problem(None, 2, mesh_text=_text("fichera.mesh"))
    m = P.mesh
    assert (m.nv, m.nedges, m.nfaces, m.ne) == (26, 51, 33, 7)
    assert P.ndofs == 26 + 51 + 33 + 7
    assert P.ndofs - P.ess.size == 7 + 9 + 3


@pytest.mark.parametrize("name", ["beam-hex.mesh", "fichera.mesh", "ref-cube.mesh"])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_file_mesh_numbering_conformity(orc, name, p):
    P = orc.Problem(None, p, mesh_text=_text(name))
    m = P.mesh
    coords = _elem_lex_coords(orc, m, P.xyz, p).reshape(-1, 3)
    lo = np.full((P.ndofs, 3), np.inf)
    hi = np.full((P.ndofs, 3), -np.inf)
    np.minimum.at(lo, P.gather_map, coords)
    np.maximum.at(hi, P.gather_map, coords)
    assert np.all(np.isfinite(lo))
    assert np.max(hi - lo) < 1e-13
    # boundary DOFs lie on the boundary of the (box or Fichera) domain
    x = lo[P.ess]
    ext = (np.isclose(x, P.xyz.min(0)) | np.isclose(x, P.xyz.max(0))).any(1)
    if name != "fichera.mesh":
        assert ext.all()
    else:  # the removed octant's three faces are boundary too
        mid = 0.5 * (P.xyz.min(0) + P.xyz.max(0))
        inner = np.isclose(x, mid).any(1)
        assert (ext | inner).all()


@pytest.mark.parametrize("name", ["beam-hex.mesh", "fichera.mesh", "ref-cube.mesh"])
def test_library_loader_matches_oracle(orc, h1pa, name):
    P = orc.Problem(None, 1, mesh_text=_text(name))
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), MESHES,
                        name)
    pm = h1pa.Mesh(path)
    m = P.mesh
    assert (pm.GetNV(), pm.GetNE(), pm.GetNEdges(), pm.GetNFaces()) == \
        (m.nv, m.ne, m.nedges, m.nfaces)
    t = pm.Topology()
    for k in ("ev", "ee", "eeo", "ef", "efo"):
        assert np.array_equal(getattr(m, k), t[k]), k
    assert np.array_equal(pm.GetVertices(), P.xyz)


@pytest.mark.parametrize("bad,code", [
    ("MFEM mesh v1.2\n", "v1.0"),
    ("MFEM mesh v1.0\ndimension\n2\n", "3D"),
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 4 0 1 2 3\n", "hexahedral"),
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 5 0 1 2 3 4 5 6 7\nboundary\n0\n"
     "vertices\n8\nnodes\n", "curved"),
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 5 0 1 2 3 4 5 6 7\nboundary\n0\n"
     "vertices\n8\n3\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n0 0 1\n1 0 1\n1 1 1\n0 1 1\n"
     .replace("0 0 1\n1 0 1\n1 1 1\n0 1 1", "0 0 -1\n1 0 -1\n1 1 -1\n0 1 -1"), "inverted"),
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 5 0 1 2 3 4 5 6 9\nboundary\n0\n"
     "vertices\n8\n3\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n0 0 1\n1 0 1\n1 1 1\n0 1 1\n", "range"),
    # 2^32 + 7 and 2^32 + 1 would wrap to valid indices if narrowed to int first
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 5 0 1 2 3 4 5 6 4294967303\nboundary\n0\n"
     "vertices\n8\n3\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n0 0 1\n1 0 1\n1 1 1\n0 1 1\n", "range"),
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 5 0 1 2 3 4 5 6 7\nboundary\n1\n"
     "1 3 0 4294967297 2 3\nvertices\n8\n3\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n0 0 1\n1 0 1\n"
     "1 1 1\n0 1 1\n", "range"),
    ("MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 5 0 1 2 3 4 5 6 7\nboundary\n1\n"
     "1 3 0 1 2 4\nvertices\n8\n3\n0 0 0\n1 0 0\n1 1 0\n0 1 0\n0 0 1\n1 0 1\n1 1 1\n"
     "0 1 1\n", "not an element face"),
])
def test_library_loader_rejects(h1pa, tmp_path, bad, code):
    f = tmp_path / "bad.mesh"
    f.write_text(bad)
    with pytest.raises(h1pa.H1paError) as ei:
        h1pa.Mesh(str(f))
    assert code in str(ei.value)


def test_file_mesh_cannot_be_perturbed(h1pa):
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), MESHES,
                        "fichera.mesh")
    pm = h1pa.Mesh(path)
    with pytest.raises(h1pa.H1paError):
        pm.Perturb(0.1, 1)


@pytest.mark.parametrize("variant", ["crlf", "comments", "tabs_in_data", "trailing"])
def test_library_loader_format_variants(h1pa, tmp_path, variant):
    """Line endings, '#' comment lines between sections (MFEM skips them),
    tabs as separators in the data and trailing blank lines read to the same
    topology."""
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = os.path.join(here, MESHES, "fichera.mesh")
    base = open(path).read()
    head, rest = base.split("\n", 1)
    txt = {"crlf": base.replace("\n", "\r\n"),
           "comments": base.replace("\nelements", "\n# comment\n\n#\nelements")
                           .replace("\nvertices", "\n# vertices follow\nvertices"),
           "tabs_in_data": head + "\n" + rest.replace(" ", "\t"),
           "trailing": base + "\n\n# end\n\n"}[variant]
    f = tmp_path / "v.mesh"
    f.write_text(txt)
    ref, got = h1pa.Mesh(path).Topology(), h1pa.Mesh(str(f)).Topology()
    for k in ("ev", "ee", "eeo", "ef", "efo"):
        assert np.array_equal(got[k], ref[k]), k
