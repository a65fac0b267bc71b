"""Pins the oracle's mesh, topology, H1 numbering and ElementRestriction.

MFEM's numbering is fixed by (a) the generalized Hilbert element order, (b)
first-encounter edge/face numbering, (c) orientation permutations.  Checks:
curve invariants, first-encounter monotonicity, MFEM's own quad orientation
table (fem/geom.cpp:969-973, the MFEM_DEBUG check of GetQuadOrientation),
geometric conformity of every shared DOF, multiplicity census and the
boundary-DOF count.  A wrong split order, a swapped edge/face table entry or a
wrong QuadDofOrd row fails at least one of them.
"""
import numpy as np
import pytest

HEX_FACES = [(3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]
HEX_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5),
             (2, 6), (3, 7)]
QUAD_ORIENT = [(0, 1, 2, 3), (0, 3, 2, 1), (1, 2, 3, 0), (1, 0, 3, 2),
               (2, 3, 0, 1), (2, 1, 0, 3), (3, 0, 1, 2), (3, 2, 1, 0)]

DIMS = [(1, 1, 1), (2, 2, 2), (3, 3, 3), (4, 4, 4), (2, 3, 4), (5, 3, 2), (3, 7, 4), (8, 8, 8)]


@pytest.mark.parametrize("dims", DIMS + [(16, 16, 16), (6, 6, 6), (9, 4, 4), (4, 9, 4), (4, 4, 9)])
def test_sfc_invariants(orc, dims):
    nx, ny, nz = dims
    c = orc.grid_sfc3d(nx, ny, nz)
    n = nx * ny * nz
    assert c.shape == (n, 3)
    lin = c[:, 0] + nx * (c[:, 1] + ny * c[:, 2])
    assert np.array_equal(np.sort(lin), np.arange(n)), "not a permutation of the cells"
    assert np.array_equal(c[0], [0, 0, 0])
    steps = np.abs(np.diff(c, axis=0)).sum(1) if n > 1 else np.zeros(0)
    if all(d % 2 == 0 or d == 1 for d in dims):
        assert np.all(steps == 1), "even grids: every step is a face neighbour"
    else:
        assert np.all(steps <= 3) and np.mean(steps == 1) > 0.8
    # the curve ends at the far end of the major axis (gilbert3d)
    major = 0 if (nx >= ny and nx >= nz) else (1 if ny >= nz else 2)
    end = [0, 0, 0]
    end[major] = dims[major] - 1
    assert np.array_equal(c[-1], end)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_sfc_power_of_two_is_hilbert(orc, k):
    """On a 2^k cube every aligned 2^m sub-cube is visited contiguously."""
    n = 2 ** k
    c = orc.grid_sfc3d(n, n, n)
    for m in range(1, k):
        blk = [tuple(v) for v in (c >> m)]
        size = 8 ** m
        for s in range(0, n ** 3, size):
            assert len(set(blk[s:s + size])) == 1


@pytest.mark.parametrize("dims", DIMS)
@pytest.mark.parametrize("sfc", [True, False])
def test_topology(orc, dims, sfc):
    nx, ny, nz = dims
    m = orc.CartesianMesh(nx, ny, nz, sfc=sfc)
    assert m.nv == (nx + 1) * (ny + 1) * (nz + 1) and m.ne == nx * ny * nz
    assert m.nedges == nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    assert m.nfaces == (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    # element vertices: unit cell of the Cartesian grid in MFEM hex order
    VTX = lambda x, y, z: x + (nx + 1) * (y + (ny + 1) * z)
    for e in range(m.ne):
        v0 = m.ev[e, 0]
        x, y, z = v0 % (nx + 1), (v0 // (nx + 1)) % (ny + 1), v0 // ((nx + 1) * (ny + 1))
        want = [VTX(x, y, z), VTX(x + 1, y, z), VTX(x + 1, y + 1, z), VTX(x, y + 1, z),
                VTX(x, y, z + 1), VTX(x + 1, y, z + 1), VTX(x + 1, y + 1, z + 1), VTX(x, y + 1, z + 1)]
        assert m.ev[e].tolist() == want
    # first-encounter numbering of edges and faces
    for tab in (m.ee, m.ef):
        seen, nxt = set(), 0
        for v in tab.ravel():
            if v not in seen:
                assert v == nxt
                seen.add(v)
                nxt += 1
    # edges: same id <=> same vertex pair; orientation is the vertex order
    pairs = {}
    for e in range(m.ne):
        for j, (a, b) in enumerate(HEX_EDGES):
            va, vb = m.ev[e, a], m.ev[e, b]
            key = (min(va, vb), max(va, vb))
            assert pairs.setdefault(key, m.ee[e, j]) == m.ee[e, j]
            assert m.eeo[e, j] == (1 if va < vb else -1)
    assert len(pairs) == m.nedges
    # faces: first element has orientation 0; the second satisfies MFEM's
    # quad Orient table: test[Orient[o][j]] == base[j]
    first = {}
    for e in range(m.ne):
        for j, fv in enumerate(HEX_FACES):
            verts = [m.ev[e, k] for k in fv]
            f = m.ef[e, j]
            if f not in first:
                first[f] = verts
                assert m.efo[e, j] == 0
            else:
                base, o = first[f], m.efo[e, j]
                assert sorted(base) == sorted(verts)
                assert all(verts[QUAD_ORIENT[o][k]] == base[k] for k in range(4))
                assert o % 2 == 1, "interior faces of a valid mesh have odd orientation"


def _elem_lex_coords(orc, m, xyz, p):
    X = m.nodes_curved(xyz, p, 0.0)  # trilinear map at the GLL nodes, lex
    D3 = (p + 1) ** 3
    return X.reshape(m.ne, 3, D3).transpose(0, 2, 1)  # [ne][D3][3]


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 4)])
def test_numbering_conformity(orc, p, dims):
    """Every E-DOF mapped to one L-DOF sits at the same physical point (on a
    perturbed mesh, so no symmetry can hide a wrong permutation)."""
    m = orc.CartesianMesh(*dims)
    xyz = m.vertices(0.2, 11)
    ed, ndofs = m.elem_dofs(p)
    nx, ny, nz = dims
    assert ndofs == (nx * p + 1) * (ny * p + 1) * (nz * p + 1)
    dm = orc.dof_map(p)
    g, off, ind = orc.restriction(ed, dm, ndofs)
    coords = _elem_lex_coords(orc, m, xyz, p).reshape(-1, 3)
    lo = np.full((ndofs, 3), np.inf)
    hi = np.full((ndofs, 3), -np.inf)
    np.minimum.at(lo, g, coords)
    np.maximum.at(hi, g, coords)
    assert np.all(np.isfinite(lo)), "every L-DOF is used"
    assert np.max(hi - lo) < 1e-13
    # native table is the lex gather map permuted by dof_map
    assert np.array_equal(ed[:, dm].ravel(), g)
    # DOF blocks: vertices, then edges, faces, bubbles (fem/fespace.cpp:3428-3535)
    assert np.array_equal(np.sort(ed[:, :8].ravel())[[0, -1]], [0, m.nv - 1])
    if p > 1:
        nb = (p - 1) ** 3
        base = m.nv + m.nedges * (p - 1) + m.nfaces * (p - 1) ** 2
        assert np.array_equal(ed[:, -nb:], base + np.arange(m.ne * nb).reshape(m.ne, nb))


@pytest.mark.parametrize("p", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("dims", [(2, 2, 2), (4, 3, 2)])
def test_restriction_tables(orc, p, dims):
    m = orc.CartesianMesh(*dims)
    ed, ndofs = m.elem_dofs(p)
    g, off, ind = orc.restriction(ed, orc.dof_map(p), ndofs)
    nE = g.size
    mult = np.diff(off)
    assert off[0] == 0 and off[-1] == nE
    assert set(np.unique(mult)) <= {1, 2, 4, 8}
    nx, ny, nz = dims
    assert np.sum(mult == 8) == (nx - 1) * (ny - 1) * (nz - 1)
    for i in range(ndofs):
        row = ind[off[i]:off[i + 1]]
        assert np.all(np.diff(row) > 0), "rows list E-indices in increasing order"
        assert np.all(g[row] == i)
    # adjointness <R x, y> = <x, R^T y>
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, ndofs)
    y = rng.uniform(-1, 1, nE)
    D3 = (p + 1) ** 3
    Rx = orc.restriction_mult(g, D3, x)
    RTy = orc.restriction_mult_transpose(off, ind, D3, y, 0)
    assert abs(Rx @ y - x @ RTy) < 1e-12 * nE
    # AddMultTranspose adds; AbsMultTranspose equals it for H1 (no signs)
    assert np.allclose(orc.restriction_mult_transpose(off, ind, D3, y, 1, y=x), x + RTy)
    assert np.array_equal(orc.restriction_mult_transpose(off, ind, D3, y, 2), RTy)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 4, 2)])
def test_boundary_dofs(orc, p, dims):
    m = orc.CartesianMesh(*dims)
    ed, ndofs = m.elem_dofs(p)
    ess = m.boundary_dofs(p, ed)
    nx, ny, nz = dims
    assert ess.size == ndofs - (nx * p - 1) * (ny * p - 1) * (nz * p - 1)
    assert np.all(np.diff(ess) > 0)
    # they are exactly the DOFs whose node lies on the cube's surface
    xyz = m.vertices()
    X = m.nodes_curved(xyz, p, 0.0).reshape(m.ne, 3, -1).transpose(0, 2, 1).reshape(-1, 3)
    g, _, _ = orc.restriction(ed, orc.dof_map(p), ndofs)
    on = np.any((np.abs(X) < 1e-12) | (np.abs(X - 1) < 1e-12), axis=1)
    assert np.array_equal(np.unique(g[on]), ess)
