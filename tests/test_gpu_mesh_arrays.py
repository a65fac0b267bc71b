"""GPU: meshes handed over as arrays (Mesh(vertices, ..., boundary, ...) +
FinalizeTopology, mesh/mesh.cpp:4891-4930), high-order nodes handed over as
the nodal space's lexicographic E-vector (GeometricFactors::Compute,
mesh/mesh.cpp:15445-15497) and ElementRestriction::BooleanMask
(fem/restriction.cpp:255-290), against the oracle on the same arrays.

Bars: restriction maps, boundary DOFs and the mask bit-exact; PA diffusion
apply and diagonal within 1e-13 relative.  One case runs through a C++
caller of include/h1pa.h (tests/capi_mesh_arrays.cpp, built by the package
build into tests/_bin).
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "_bin", "capi_mesh_arrays")
RTOL = 1e-13


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _fichera(orc):
    with open(os.path.join(ROOT, "tests", "golden", "meshes", "fichera.mesh")) as f:
        return orc.read_mfem_mesh(f.read())


def _shuffled_grid(orc, nx, ny, nz, seed):
    """A Cartesian grid in lexicographic order with its elements shuffled
    and its interior vertices jittered: a mesh no Cartesian constructor
    makes, given only as arrays."""
    m = orc.CartesianMesh(nx, ny, nz, sfc=False)
    xyz = m.vertices(0.15, seed)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(m.ne)
    ev = m.ev[perm].copy()
    # boundary quads: the element faces on the cube's faces
    faces = [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]
    bq = []
    for e in range(m.ne):
        for f in faces:
            vv = [int(ev[e, k]) for k in f]
            c = xyz[vv]
            for ax in range(3):
                if np.all(np.abs(c[:, ax]) < 1e-12) or np.all(np.abs(c[:, ax] - 1) < 1e-12):
                    bq.append(vv)
                    break
    return ev, np.array(bq, np.int32), xyz


def _case(orc, name):
    if name == "fichera":
        return _fichera(orc)
    return _shuffled_grid(orc, 3, 4, 2, 11)


@pytest.mark.parametrize("name", ["fichera", "grid"])
@pytest.mark.parametrize("p", [1, 3, 5])
def test_arrays_mesh_matches_oracle(orc, h1pa, torch_cuda, name, p):
    M, torch = h1pa, torch_cuda
    print(f"[case] arrays mesh={name} p={p}", flush=True)
    ev, bq, xyz = _case(orc, name)
    mesh = M.Mesh.FromArrays(xyz, ev, bq)
    assert mesh.GetNE() == ev.shape[0] and mesh.GetNV() == xyz.shape[0]
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    P = orc.Problem(None, p, mesh_arrays=(ev, bq, xyz), coeff="sincos")
    assert fes.GetNDofs() == P.ndofs
    R = fes.GetElementRestriction()
    assert np.array_equal(R.GatherMap().cpu().numpy(), P.gather_map)
    assert np.array_equal(R.Offsets().cpu().numpy(), P.offsets)
    assert np.array_equal(R.Indices().cpu().numpy(), P.indices)
    assert np.array_equal(fes.GetBoundaryTrueDofs().cpu().numpy(), P.ess)
    mask = torch.empty(P.ne * P.D ** 3, dtype=torch.float64, device="cuda")
    R.BooleanMask(mask)
    assert np.array_equal(mask.cpu().numpy(), orc.boolean_mask(P.offsets, P.indices, mask.numel()))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient()))
    a.Assemble()
    x = np.random.default_rng(p).uniform(-1, 1, P.ndofs)
    y = torch.empty(P.ndofs, dtype=torch.float64, device="cuda")
    a.Mult(torch.from_numpy(x).cuda(), y)
    assert _rel(y.cpu().numpy(), P.mult(x)) < RTOL
    d = torch.empty_like(y)
    a.AssembleDiagonal(d)
    assert _rel(d.cpu().numpy(), P.diagonal()) < RTOL


@pytest.mark.parametrize("pg,p", [(2, 2), (3, 4)])
def test_given_nodes_match_oracle(orc, h1pa, torch_cuda, pg, p):
    """Curved order-pg nodes handed over as an E-vector give the oracle's
    operator on those nodes; the same nodes made by SetCurvature give
    bitwise the same apply."""
    M, torch = h1pa, torch_cuda
    print(f"[case] given nodes pg={pg} p={p}", flush=True)
    ev, bq, xyz = _shuffled_grid(orc, 3, 3, 3, 5)
    om = orc.CartesianMesh(0, 0, 0, elements=(ev, bq, xyz.shape[0]))
    X = om.nodes_curved(xyz, pg, 0.05)
    mesh = M.Mesh.FromArrays(xyz, ev, bq)
    mesh.SetNodes(pg, torch.from_numpy(X).cuda())
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    P = orc.Problem(None, p, mesh_arrays=(ev, bq, xyz), nodes=(pg, X), coeff="sincos")
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient()))
    a.Assemble()
    x = np.random.default_rng(pg).uniform(-1, 1, P.ndofs)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    a.Mult(xd, y)
    assert _rel(y.cpu().numpy(), P.mult(x)) < RTOL
    # the library's own curving map on the same mesh: the same nodes
    mesh2 = M.Mesh.FromArrays(xyz, ev, bq)
    mesh2.SetCurvature(pg, 0.05)
    fes2 = M.FiniteElementSpace(mesh2, M.H1_FECollection(p, 3))
    a2 = M.BilinearForm(fes2)
    a2.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a2.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient()))
    a2.Assemble()
    y2 = torch.empty_like(xd)
    a2.Mult(xd, y2)
    assert _rel(y2.cpu().numpy(), y.cpu().numpy()) < RTOL


def test_bad_arrays_rejected(h1pa, torch_cuda):
    M = h1pa
    ev = np.arange(8, dtype=np.int32).reshape(1, 8)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
    with pytest.raises(M.H1paError):
        M.Mesh.FromArrays(xyz, ev + 1, np.zeros((0, 4), np.int32))   # index 8 out of range
    with pytest.raises(M.H1paError):
        M.Mesh.FromArrays(xyz, ev, np.array([[0, 1, 6, 7]], np.int32))  # not a face
    with pytest.raises(M.H1paError):
        M.Mesh.FromArrays(xyz[[4, 5, 6, 7, 0, 1, 2, 3]].copy(), ev,
                          np.zeros((0, 4), np.int32))  # z flipped: inverted
    m = M.Mesh.FromArrays(xyz, ev, np.array([[0, 3, 2, 1]], np.int32))
    assert m.GetNE() == 1 and m.GetNFaces() == 6 and m.GetNEdges() == 12


@pytest.mark.parametrize("name,p,pg", [("fichera", 3, 0), ("grid", 2, 3)])
def test_cpp_caller_matches_oracle(orc, h1pa, torch_cuda, tmp_path, name, p, pg):
    """The C++ caller builds the mesh from arrays (and nodes) through the C
    ABI; its maps, boundary DOFs and mask are bit-exact, its apply and
    diagonal within 1e-13 of the oracle on the same arrays."""
    print(f"[case] C++ caller mesh={name} p={p} pg={pg}", flush=True)
    assert os.path.exists(EXE), "tests/_bin/capi_mesh_arrays is not built"
    ev, bq, xyz = _case(orc, name)
    nodes = None
    if pg:
        om = orc.CartesianMesh(0, 0, 0, elements=(ev, bq, xyz.shape[0]))
        nodes = (pg, om.nodes_curved(xyz, pg, 0.05))
    P = orc.Problem(None, p, mesh_arrays=(ev, bq, xyz), nodes=nodes, coeff="sincos")
    x = np.random.default_rng(3).uniform(-1, 1, P.ndofs)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([xyz.shape[0], ev.shape[0], bq.shape[0], p, pg, P.ndofs], np.int64).tofile(f)
        np.ascontiguousarray(xyz, np.float64).tofile(f)
        np.ascontiguousarray(ev, np.int32).tofile(f)
        np.ascontiguousarray(bq, np.int32).tofile(f)
        if pg:
            np.ascontiguousarray(nodes[1], np.float64).tofile(f)
        x.tofile(f)
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    ndofs, dof, n_ess = np.frombuffer(raw[:24], np.int64)
    nE = P.ne * dof
    sizes = [(np.int32, nE), (np.int32, ndofs + 1), (np.int32, nE), (np.int32, n_ess),
             (np.float64, ndofs), (np.float64, ndofs), (np.float64, nE)]
    out, off = [], 24
    for dt, n in sizes:
        a = np.frombuffer(raw, dt, count=int(n), offset=off)
        out.append(a)
        off += a.nbytes
    assert off == len(raw)
    g, o, i, ess, y, diag, mask = out
    assert ndofs == P.ndofs
    assert np.array_equal(g, P.gather_map)
    assert np.array_equal(o, P.offsets)
    assert np.array_equal(i, P.indices)
    assert np.array_equal(ess, P.ess)
    assert np.array_equal(mask, orc.boolean_mask(P.offsets, P.indices, nE))
    assert _rel(y, P.mult(x)) < RTOL
    assert _rel(diag, P.diagonal()) < RTOL
