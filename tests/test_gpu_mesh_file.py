"""GPU: spaces, PA operators and solves on meshes read from MFEM v1.0 files,
against the oracle on the same file.

Bars: restriction maps, boundary DOFs and the gather / scatter operations
bit-exact; PA diffusion and mass
applies within 1e-13 relative; Jacobi-PCG iteration count within +-1 and
solution within 1e-9 (rel_tol 1e-12).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAMES = ["beam-hex.mesh", "fichera.mesh", "ref-cube.mesh"]


def _path(name):
    return os.path.join(HERE, "tests", "golden", "meshes", name)


class FilePair:
    def __init__(self, orc, M, torch, name, p, coeff="sincos", integrators=("diffusion",),
                 curve=0.0):
        print(f"[case] file={name} p={p} coeff={coeff} integ={integrators} curve={curve}",
              flush=True)
        self.torch = torch
        self.mesh = M.Mesh(_path(name))
        if curve:
            self.mesh.SetCurvature(p, curve)
        self.fes = M.FiniteElementSpace(self.mesh, M.H1_FECollection(p, 3))
        self.a = M.BilinearForm(self.fes)
        self.a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
        coef = M.SinCosCoefficient() if coeff == "sincos" else M.ConstantCoefficient(1.0)
        for it in integrators:
            cls = M.DiffusionIntegrator if it == "diffusion" else M.MassIntegrator
            self.a.AddDomainIntegrator(cls(coef))
        self.a.Assemble()
        with open(_path(name)) as f:
            self.P = orc.Problem(None, p, mesh_text=f.read(), coeff=coeff,
                                 integrators=integrators, curve=curve)
        assert self.fes.GetNDofs() == self.P.ndofs

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def host(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("p", [1, 2, 4, 7])
def test_maps_and_boundary_bit_exact(orc, h1pa, torch_cuda, name, p):
    S = FilePair(orc, h1pa, torch_cuda, name, p)
    R = S.fes.GetElementRestriction()
    g, off, ind = R.GatherMap(), R.Offsets(), R.Indices()
    assert np.array_equal(S.host(g), S.P.gather_map)
    assert np.array_equal(S.host(off), S.P.offsets)
    assert np.array_equal(S.host(ind), S.P.indices)
    assert np.array_equal(S.host(S.fes.GetBoundaryTrueDofs()), S.P.ess)
    # gather and the transpose family on the file mesh's own numbering
    D3 = (p + 1) ** 3
    rng = np.random.default_rng(p)
    x = rng.uniform(-1, 1, S.P.ndofs)
    xe = rng.uniform(-1, 1, S.P.gather_map.size)
    ye = S.torch.empty(xe.size, dtype=S.torch.float64, device="cuda")
    R.Mult(S.dev(x), ye)
    assert np.array_equal(S.host(ye), orc.restriction_mult(S.P.gather_map, D3, x))
    y = S.torch.empty(S.P.ndofs, dtype=S.torch.float64, device="cuda")
    R.MultTranspose(S.dev(xe), y)
    assert np.array_equal(S.host(y), orc.restriction_mult_transpose(S.P.offsets, S.P.indices, D3,
                                                                     xe, 0))
    R.AbsMultTranspose(S.dev(xe), y)
    assert np.array_equal(S.host(y), orc.restriction_mult_transpose(S.P.offsets, S.P.indices, D3,
                                                                     xe, 2))


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("p,integ", [(1, ("diffusion",)), (3, ("diffusion", "mass")),
                                     (5, ("diffusion",)), (8, ("mass",))])
def test_apply_matches_oracle(orc, h1pa, torch_cuda, name, p, integ):
    S = FilePair(orc, h1pa, torch_cuda, name, p, integrators=integ)
    x = np.random.default_rng(p).uniform(-1, 1, S.P.ndofs)
    y = torch_cuda.empty(S.P.ndofs, dtype=torch_cuda.float64, device="cuda")
    S.a.Mult(S.dev(x), y)
    err = _rel(S.host(y), S.P.mult(x))
    print(f"[file] {name} p={p} {integ} apply err={err:.2e}", flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("name,p,curve", [("fichera.mesh", 3, 0.0), ("beam-hex.mesh", 4, 0.0),
                                          ("fichera.mesh", 2, 0.03)])
def test_pcg_on_file_mesh(orc, h1pa, torch_cuda, name, p, curve):
    torch = torch_cuda
    S = FilePair(orc, h1pa, torch, name, p, curve=curve)
    b = np.random.default_rng(7).uniform(-1, 1, S.P.ndofs)
    ess = S.fes.GetBoundaryTrueDofs()
    A, X, B = S.a.FormLinearSystem(ess, torch.zeros(S.P.ndofs, dtype=torch.float64,
                                                    device="cuda"), S.dev(b))
    cg = h1pa.CGSolver()
    cg.SetRelTol(1e-12)
    cg.SetMaxIter(2000)
    cg.SetOperator(A)
    cg.SetPreconditioner(h1pa.OperatorJacobiSmoother(S.a, ess))
    cg.Mult(B, X)
    _, Bo = S.P.form_linear_system(np.zeros(S.P.ndofs), b)
    xo, st = S.P.pcg(Bo, rel_tol=1e-12, max_iter=2000)
    print(f"[file] {name} p={p} pcg iters gpu={cg.GetNumIterations()} oracle={st['iterations']}",
          flush=True)
    assert cg.GetConverged() and st["converged"]
    assert abs(cg.GetNumIterations() - st["iterations"]) <= 1
    assert _rel(S.host(X), xo) < 1e-9
