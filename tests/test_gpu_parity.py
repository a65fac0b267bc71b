"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: integer maps and restriction sums bit-exact; FP64 operator results within
1e-13 relative (norm-wise); PCG iteration counts within +-1.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-13


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


class Pair:
    """The same problem built by the product and by the oracle."""

    def __init__(self, orc, M, torch, dims, p, q1d=None, perturb=0.0, seed=0, curve=0.0,
                 coeff="one", integrators=("diffusion",)):
        print(f"[case] dims={dims} p={p} q1d={q1d} perturb={perturb} curve={curve} "
              f"coeff={coeff} integ={integrators}", flush=True)
        self.torch = torch
        self.mesh = M.Mesh.MakeCartesian3D(*dims)
        if perturb:
            self.mesh.Perturb(perturb, seed)
        if curve:
            self.mesh.SetCurvature(p, curve)
        self.fes = M.FiniteElementSpace(self.mesh, M.H1_FECollection(p, 3))
        self.a = M.BilinearForm(self.fes)
        self.a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
        coef = M.SinCosCoefficient() if coeff == "sincos" else M.ConstantCoefficient(1.0)
        for it in integrators:
            cls = M.DiffusionIntegrator if it == "diffusion" else M.MassIntegrator
            self.a.AddDomainIntegrator(cls(coef, q1d or 0))
        self.a.Assemble()
        self.P = orc.Problem(None, p, Q=q1d, perturb=perturb, seed=seed, curve=curve, coeff=coeff,
                             integrators=integrators, dims=dims)
        assert self.fes.GetNDofs() == self.P.ndofs

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def host(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()


DIMS = [(1, 1, 1), (2, 2, 2), (3, 2, 4), (4, 4, 4)]


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("dims", DIMS)
def test_restriction_maps_bit_exact(orc, h1pa, torch_cuda, p, dims):
    if p >= 6 and dims == (4, 4, 4):
        dims = (3, 3, 3)
    print(f"[case] maps dims={dims} p={p}", flush=True)
    m = h1pa.Mesh.MakeCartesian3D(*dims)
    fes = h1pa.FiniteElementSpace(m, h1pa.H1_FECollection(p))
    om = orc.CartesianMesh(*dims)
    ed, ndofs = om.elem_dofs(p)
    g, off, ind = orc.restriction(ed, orc.dof_map(p), ndofs)
    R = fes.GetElementRestriction()
    assert fes.GetNDofs() == ndofs
    assert np.array_equal(R.GatherMap().cpu().numpy(), g)
    assert np.array_equal(R.Offsets().cpu().numpy(), off)
    assert np.array_equal(R.Indices().cpu().numpy(), ind)
    assert np.array_equal(fes.GetElementToDofTable().cpu().numpy(), ed)
    assert np.array_equal(fes.GetBoundaryTrueDofs().cpu().numpy(), om.boundary_dofs(p, ed))


@pytest.mark.parametrize("p", [1, 2, 3, 5, 8])
def test_restriction_ops_bit_exact(orc, h1pa, torch_cuda, p):
    torch = torch_cuda
    dims = (3, 2, 4) if p < 8 else (2, 2, 2)
    print(f"[case] restriction ops dims={dims} p={p}", flush=True)
    m = h1pa.Mesh.MakeCartesian3D(*dims)
    fes = h1pa.FiniteElementSpace(m, h1pa.H1_FECollection(p))
    R = fes.GetElementRestriction()
    om = orc.CartesianMesh(*dims)
    ed, ndofs = om.elem_dofs(p)
    g, off, ind = orc.restriction(ed, orc.dof_map(p), ndofs)
    D3 = (p + 1) ** 3
    rng = np.random.default_rng(p)
    x = rng.uniform(-1, 1, ndofs)
    xe = rng.uniform(-1, 1, g.size)
    y0 = rng.uniform(-1, 1, ndofs)
    dx, dxe = torch.from_numpy(x).cuda(), torch.from_numpy(xe).cuda()
    ye = torch.empty(g.size, dtype=torch.float64, device="cuda")
    R.Mult(dx, ye)
    assert np.array_equal(ye.cpu().numpy(), orc.restriction_mult(g, D3, x))
    # an output only 8-byte aligned (a view one double in) takes the
    # entry-by-entry path: same bits
    buf = torch.zeros(g.size + 1, dtype=torch.float64, device="cuda")
    R.Mult(dx, buf[1:])
    assert np.array_equal(buf[1:].cpu().numpy(), orc.restriction_mult(g, D3, x))
    R.AbsMult(dx, ye)
    assert np.array_equal(ye.cpu().numpy(), orc.restriction_mult(g, D3, x))
    y = torch.empty(ndofs, dtype=torch.float64, device="cuda")
    R.MultTranspose(dxe, y)
    assert np.array_equal(y.cpu().numpy(), orc.restriction_mult_transpose(off, ind, D3, xe, 0))
    y = torch.from_numpy(y0.copy()).cuda()
    R.AddMultTranspose(dxe, y)
    assert np.array_equal(y.cpu().numpy(), orc.restriction_mult_transpose(off, ind, D3, xe, 1, y=y0))
    R.AbsMultTranspose(dxe, y)
    assert np.array_equal(y.cpu().numpy(), orc.restriction_mult_transpose(off, ind, D3, xe, 2))
    R.MultLeftInverse(dxe, y)
    assert np.array_equal(y.cpu().numpy(), xe[ind[off[1:] - 1]])
    with pytest.raises(h1pa.H1paError):
        R.AddMultTranspose(dxe, y, 2.0)
    # deterministic: repeated scatters are bitwise identical
    y2 = torch.empty_like(y)
    R.MultTranspose(dxe, y)
    R.MultTranspose(dxe, y2)
    assert torch.equal(y, y2)


@pytest.mark.parametrize("p", [1, 3, 6])
def test_restriction_from_mfem_table(orc, h1pa, torch_cuda, p):
    """ElementRestriction built from MFEM's own element-to-DOF table and
    dof_map, including signed (-1-i) entries (restriction.cpp:26-114)."""
    om = orc.CartesianMesh(3, 2, 2)
    ed, ndofs = om.elem_dofs(p)
    dm = orc.dof_map(p)
    rng = np.random.default_rng(0)
    ed_s = ed.copy()
    flip = rng.random(ed.shape) < 0.2
    ed_s[flip] = -1 - ed_s[flip]
    for table in (ed, ed_s):
        print(f"[case] from-table p={p} signed={table is ed_s}", flush=True)
        g, off, ind = orc.restriction(table, dm, ndofs)
        R = h1pa.ElementRestriction.FromTable(table, dm, ndofs)
        assert np.array_equal(R.GatherMap().cpu().numpy(), g)
        assert np.array_equal(R.Offsets().cpu().numpy(), off)
        assert np.array_equal(R.Indices().cpu().numpy(), ind)
        x = rng.uniform(-1, 1, ndofs)
        xe = rng.uniform(-1, 1, g.size)
        D3 = (p + 1) ** 3
        ye = torch_cuda.empty(g.size, dtype=torch_cuda.float64, device="cuda")
        R.Mult(torch_cuda.from_numpy(x).cuda(), ye)
        assert np.array_equal(ye.cpu().numpy(), orc.restriction_mult(g, D3, x))
        y = torch_cuda.empty(ndofs, dtype=torch_cuda.float64, device="cuda")
        R.MultTranspose(torch_cuda.from_numpy(xe).cuda(), y)
        assert np.array_equal(y.cpu().numpy(), orc.restriction_mult_transpose(off, ind, D3, xe, 0))


@pytest.mark.parametrize("p", [2, 5])
def test_integrators_on_different_rules(orc, h1pa, torch_cuda, p):
    """Diffusion on Q = p+2 and mass on Q = p+1 in one BilinearForm (two
    operators summed on the device): apply and diagonal within 1e-13 of the
    oracle's two operators added."""
    torch, M = torch_cuda, h1pa
    dims = (3, 2, 2)
    mesh = M.Mesh.MakeCartesian3D(*dims)
    mesh.Perturb(0.2, 9)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient(), p + 2))
    a.AddDomainIntegrator(M.MassIntegrator(M.SinCosCoefficient(), p + 1))
    a.Assemble()
    Pd = orc.Problem(None, p, Q=p + 2, perturb=0.2, seed=9, coeff="sincos",
                     integrators=("diffusion",), dims=dims)
    Pm = orc.Problem(None, p, Q=p + 1, perturb=0.2, seed=9, coeff="sincos",
                     integrators=("mass",), dims=dims)
    x = np.random.default_rng(p).uniform(-1, 1, Pd.ndofs)
    y = torch.empty(Pd.ndofs, dtype=torch.float64, device="cuda")
    a.Mult(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy(), Pd.mult(x) + Pm.mult(x)) < 1e-13
    d = torch.empty_like(y)
    a.AssembleDiagonal(d)
    torch.cuda.synchronize()
    assert _rel(d.cpu().numpy(), Pd.diagonal() + Pm.diagonal()) < 1e-13


@pytest.mark.parametrize("p", [2, 4, 7])
def test_add_mult_and_transpose(orc, h1pa, torch_cuda, p):
    """Operator::AddMult (y += a A x, linalg/operator.cpp:51-56) within 1e-13 of
    the oracle; MultTranspose = Mult bitwise (symmetric operators);
    AddMultTranspose = AddMult bitwise."""
    torch = torch_cuda
    S = Pair(orc, h1pa, torch, (3, 2, 2), p, perturb=0.2, seed=p, coeff="sincos",
             integrators=("diffusion", "mass"))
    rng = np.random.default_rng(p)
    x = rng.uniform(-1, 1, S.P.ndofs)
    y0 = rng.uniform(-1, 1, S.P.ndofs)
    y = S.dev(y0.copy())
    S.a.AddMult(S.dev(x), y, -0.7)
    ref = y0 - 0.7 * S.P.mult(x)
    assert _rel(S.host(y), ref) < 1e-13
    y2 = S.dev(y0.copy())
    S.a.AddMultTranspose(S.dev(x), y2, -0.7)
    assert torch.equal(y, y2)
    m1 = torch.empty_like(y)
    m2 = torch.empty_like(y)
    S.a.Mult(S.dev(x), m1)
    S.a.MultTranspose(S.dev(x), m2)
    assert torch.equal(m1, m2)


@pytest.mark.parametrize("p", [1, 3, 5])
def test_native_ordering_restriction(orc, h1pa, torch_cuda, p):
    """GetElementRestriction(NATIVE): the element-to-DOF table without the
    lexicographic reordering (restriction.cpp:60-98), maps and operations
    bit-exact against the oracle built with no dof_map."""
    torch = torch_cuda
    dims = (3, 2, 2)
    m = h1pa.Mesh.MakeCartesian3D(*dims)
    fes = h1pa.FiniteElementSpace(m, h1pa.H1_FECollection(p))
    R = fes.GetElementRestriction(h1pa.ElementDofOrdering.NATIVE)
    assert fes.GetElementRestriction(h1pa.ElementDofOrdering.NATIVE) is R
    om = orc.CartesianMesh(*dims)
    ed, ndofs = om.elem_dofs(p)
    g, off, ind = orc.restriction(ed, None, ndofs)
    assert np.array_equal(R.GatherMap().cpu().numpy(), g)
    assert np.array_equal(R.Offsets().cpu().numpy(), off)
    assert np.array_equal(R.Indices().cpu().numpy(), ind)
    D3 = (p + 1) ** 3
    rng = np.random.default_rng(p)
    x = rng.uniform(-1, 1, ndofs)
    xe = rng.uniform(-1, 1, g.size)
    ye = torch.empty(g.size, dtype=torch.float64, device="cuda")
    R.Mult(torch.from_numpy(x).cuda(), ye)
    assert np.array_equal(ye.cpu().numpy(), orc.restriction_mult(g, D3, x))
    y = torch.empty(ndofs, dtype=torch.float64, device="cuda")
    R.MultTranspose(torch.from_numpy(xe).cuda(), y)
    assert np.array_equal(y.cpu().numpy(), orc.restriction_mult_transpose(off, ind, D3, xe, 0))
    with pytest.raises(h1pa.H1paError):
        fes.GetElementRestriction(7)


@pytest.mark.parametrize("case", [dict(perturb=0.2, seed=3), dict(curve=0.05)])
@pytest.mark.parametrize("p", [1, 2, 4, 6])
def test_geometric_factors(orc, h1pa, torch_cuda, p, case):
    S = Pair(orc, h1pa, torch_cuda, (3, 2, 2), p, **case)
    X, J, dJ = S.mesh.GetGeometricFactors(p + 2)
    assert _rel(S.host(X), S.P.P) < RTOL
    assert _rel(S.host(J), S.P.J) < RTOL
    Jr = S.P.J.reshape(S.P.ne, 3, 3, -1)
    det_ref = np.linalg.det(Jr.transpose(0, 3, 2, 1))  # [e][q][r][c]
    assert _rel(S.host(dJ), det_ref.ravel()) < RTOL


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("integ", [("diffusion",), ("mass",), ("diffusion", "mass")])
def test_pa_data_and_apply(orc, h1pa, torch_cuda, p, integ):
    dims = (3, 2, 2) if p <= 5 else (2, 2, 1)
    S = Pair(orc, h1pa, torch_cuda, dims, p, perturb=0.2, seed=p, coeff="sincos",
             integrators=integ)
    dd, dm = S.a.PAData()
    if "diffusion" in integ:
        assert _rel(S.host(dd), S.P.Dd) < RTOL
    if integ == ("mass",):
        assert _rel(S.host(dm), S.P.Mv) < RTOL
    rng = np.random.default_rng(10 + p)
    x = rng.uniform(-1, 1, S.P.ndofs)
    y = torch_cuda.empty(S.P.ndofs, dtype=torch_cuda.float64, device="cuda")
    S.a.Mult(S.dev(x), y)
    assert _rel(S.host(y), S.P.mult(x)) < RTOL
    d = torch_cuda.empty_like(y)
    S.a.AssembleDiagonal(d)
    assert _rel(S.host(d), S.P.diagonal()) < RTOL


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6, 7, 8])
def test_apply_q_p_plus_1(orc, h1pa, torch_cuda, p):
    """q1d = p+1 (the other compiled quadrature) on a curved mesh."""
    S = Pair(orc, h1pa, torch_cuda, (2, 2, 1), p, q1d=p + 1, curve=0.05,
             integrators=("diffusion", "mass"))
    x = np.random.default_rng(p).uniform(-1, 1, S.P.ndofs)
    y = torch_cuda.empty(S.P.ndofs, dtype=torch_cuda.float64, device="cuda")
    S.a.Mult(S.dev(x), y)
    assert _rel(S.host(y), S.P.mult(x)) < RTOL


@pytest.mark.parametrize("p", [2, 5])
def test_add_mult_e(orc, h1pa, torch_cuda, p):
    S = Pair(orc, h1pa, torch_cuda, (2, 3, 2), p, perturb=0.2, seed=1)
    nE = S.P.gather_map.size
    rng = np.random.default_rng(2)
    xe, ye0 = rng.uniform(-1, 1, nE), rng.uniform(-1, 1, nE)
    ye = S.dev(ye0.copy())
    S.a.AddMultPA_E(S.dev(xe), ye)
    ref = orc.diffusion_apply(p, S.P.Q, S.P.ne, S.P.B, S.P.G, S.P.Dd, xe, ye0)
    assert _rel(S.host(ye), ref) < RTOL


@pytest.mark.parametrize("p", [1, 3, 4])
def test_constrained_system_and_rhs(orc, h1pa, torch_cuda, p):
    torch = torch_cuda
    S = Pair(orc, h1pa, torch, (3, 3, 2), p, perturb=0.1, seed=7)
    M = h1pa
    ess = S.fes.GetBoundaryTrueDofs()
    lf = M.LinearForm(S.fes)
    lf.AddDomainIntegrator(M.DomainLFIntegrator(M.ConstantCoefficient(1.0)))
    b = lf.Assemble()
    assert _rel(S.host(b), S.P.rhs(1.0)) < RTOL
    rng = np.random.default_rng(3)
    x0 = rng.uniform(-1, 1, S.P.ndofs)  # non-zero boundary values
    A, X, B = S.a.FormLinearSystem(ess, S.dev(x0), b)
    Xr, Br = S.P.form_linear_system(x0, S.P.rhs(1.0))
    assert np.array_equal(S.host(X), Xr)
    assert _rel(S.host(B), Br) < RTOL
    v = rng.uniform(-1, 1, S.P.ndofs)
    y = torch.empty(S.P.ndofs, dtype=torch.float64, device="cuda")
    A.Mult(S.dev(v), y)
    assert _rel(S.host(y), S.P.constrained_mult(v)) < RTOL


@pytest.mark.parametrize("p", [2, 5])
def test_rhs_quadrature_function(orc, h1pa, torch_cuda, p):
    """DomainLFIntegrator(QuadratureFunctionCoefficient) on a curved mesh: f
    evaluated at the points the product reports (geometric factors at
    q1d = p + 1), the same values handed to the oracle."""
    torch = torch_cuda
    M = h1pa
    S = Pair(orc, h1pa, torch, (3, 2, 3), p, curve=0.05)
    Q = p + 1
    nq, ne = Q ** 3, S.P.ne
    X, _, _ = S.mesh.GetGeometricFactors(Q, M.QI_VALUES)
    Xh = S.host(X).reshape(ne, 3, nq)
    x, y, z = Xh[:, 0], Xh[:, 1], Xh[:, 2]
    fq = (1.0 + x * y - z * z + np.sin(3 * x)).ravel()  # [ne][nq]
    lf = M.LinearForm(S.fes)
    lf.AddDomainIntegrator(M.DomainLFIntegrator(M.QuadratureFunctionCoefficient(S.dev(fq))))
    b = lf.Assemble()
    assert _rel(S.host(b), S.P.rhs(fq=fq)) < RTOL
    # a constant quadrature function is the constant source
    lf = M.LinearForm(S.fes)
    lf.AddDomainIntegrator(M.DomainLFIntegrator(
        M.QuadratureFunctionCoefficient(S.dev(np.full(nq * ne, 1.5)))))
    b = lf.Assemble()
    assert _rel(S.host(b), S.P.rhs(1.5)) < RTOL
    # a quadrature function of the wrong size is refused
    lf = M.LinearForm(S.fes)
    lf.AddDomainIntegrator(M.DomainLFIntegrator(M.QuadratureFunctionCoefficient(S.dev(fq[:-1]))))
    with pytest.raises(ValueError, match="QuadratureFunctionCoefficient"):
        lf.Assemble()


def test_pcg_cfg1(orc, h1pa, torch_cuda):
    """Config 1: 16^3, p = 3, k = 1, f = 1, homogeneous Dirichlet, Jacobi-PCG
    with CGSolver::SetRelTol(1e-12) semantics (preconditioned norm)."""
    torch = torch_cuda
    M = h1pa
    S = Pair(orc, h1pa, torch, (16, 16, 16), 3)
    assert S.P.ndofs == 117649
    ess = S.fes.GetBoundaryTrueDofs()
    assert ess.numel() == 117649 - 47 ** 3
    lf = M.LinearForm(S.fes)
    lf.AddDomainIntegrator(M.DomainLFIntegrator(M.ConstantCoefficient(1.0)))
    b = lf.Assemble()
    x = torch.zeros(S.P.ndofs, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, x, b)
    cg = M.CGSolver()
    cg.SetRelTol(1e-12)
    cg.SetMaxIter(2000)
    cg.SetOperator(A)
    cg.SetPreconditioner(M.OperatorJacobiSmoother(S.a, ess))
    cg.Mult(B, X)
    assert cg.GetConverged()
    bo = S.P.rhs(1.0)
    Xo, Bo = S.P.form_linear_system(np.zeros(S.P.ndofs), bo)
    xo, st = S.P.pcg(Bo, rel_tol=1e-12, max_iter=2000)
    print(f"[pcg] gpu iters={cg.GetNumIterations()} oracle iters={st['iterations']}", flush=True)
    assert abs(cg.GetNumIterations() - st["iterations"]) <= 1
    assert _rel(S.host(X), xo) < 1e-9
    # true residual of the constrained system
    r = Bo - S.P.constrained_mult(S.host(X))
    assert np.linalg.norm(r) / np.linalg.norm(Bo) < 1e-9


def test_large_size_properties(h1pa, torch_cuda):
    """~4M-DOF p = 4 operator (perturbed, variable k): A 1 = 0, symmetry,
    exact energy of a linear function, determinism."""
    torch = torch_cuda
    M = h1pa
    n, p = 40, 4
    print(f"[case] large n={n} p={p}", flush=True)
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, 99)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.ConstantCoefficient(1.0)))
    a.Assemble()
    N = fes.GetNDofs()
    assert N == (n * p + 1) ** 3
    one = torch.ones(N, dtype=torch.float64, device="cuda")
    y = torch.empty_like(one)
    a.Mult(one, y)
    gen = torch.Generator(device="cuda").manual_seed(0)
    u = torch.rand(N, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    v = torch.rand(N, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    Au, Av = torch.empty_like(u), torch.empty_like(u)
    a.Mult(u, Au)
    a.Mult(v, Av)
    scale = Au.abs().max().item()
    assert y.abs().max().item() < 1e-12 * scale
    s1, s2 = torch.dot(v, Au).item(), torch.dot(u, Av).item()
    assert abs(s1 - s2) < 1e-11 * abs(s1)
    Au2 = torch.empty_like(u)
    a.Mult(u, Au2)
    assert torch.equal(Au, Au2)
    # linear function u = x + 2y + 3z at the DOF nodes (trilinear map of each
    # element at the GLL nodes) -> a(u, u) = |(1,2,3)|^2 |Omega| = 14
    import oracle
    g = fes.GetElementRestriction().GatherMap().long()
    xyz = torch.from_numpy(mesh.GetVertices()).cuda()
    topo = mesh.Topology()
    ev = torch.from_numpy(topo["ev"]).cuda().long()
    zl = torch.tensor(oracle.gauss_lobatto(p + 1)[0], dtype=torch.float64, device="cuda")
    D = p + 1
    s = torch.stack(torch.meshgrid(zl, zl, zl, indexing="ij"), -1).permute(2, 1, 0, 3).reshape(-1, 3)
    lex2v = [0, 1, 3, 2, 4, 5, 7, 6]
    V = xyz[ev[:, lex2v]]  # [ne][8][3]
    Nw = []
    for l in range(8):
        w = torch.ones(s.shape[0], dtype=torch.float64, device="cuda")
        for c in range(3):
            w = w * (s[:, c] if (l >> c) & 1 else 1 - s[:, c])
        Nw.append(w)
    Nw = torch.stack(Nw, 1)  # [D3][8]
    Xe = torch.einsum("dl,elc->edc", Nw, V)  # [ne][D3][3]
    ue = Xe[..., 0] + 2 * Xe[..., 1] + 3 * Xe[..., 2]
    ul = torch.zeros(N, dtype=torch.float64, device="cuda")
    ul[g] = ue.reshape(-1)
    Aul = torch.empty_like(ul)
    a.Mult(ul, Aul)
    assert abs(torch.dot(ul, Aul).item() - 14.0) < 1e-10
    assert D ** 3 == fes.dof


@pytest.mark.parametrize("p", [2, 5])
def test_quadrature_function_coefficient(orc, h1pa, torch_cuda, p):
    """QuadratureFunctionCoefficient (MFEM's general coefficient path:
    CoefficientVector::Project of a QuadratureFunction, fem/coefficient.cpp:
    2058-2076): the sincos field evaluated on the host side at the mesh's own
    physical quadrature points, q fastest per element, gives the SinCos
    form's PA data and the oracle's apply; a wrongly sized array is refused."""
    M, torch = h1pa, torch_cuda
    dims = (3, 2, 2)
    S = Pair(orc, M, torch, dims, p, perturb=0.2, seed=p, coeff="sincos",
             integrators=("diffusion", "mass"))
    Q, ne = p + 2, S.fes.GetNE()
    X, _, _ = S.mesh.GetGeometricFactors(Q, M.QI_VALUES)
    Xr = X.view(ne, 3, Q ** 3)
    tp = 2 * np.pi
    k = (1 + 0.5 * torch.sin(tp * Xr[:, 0]) * torch.cos(tp * Xr[:, 1]) *
         torch.sin(tp * Xr[:, 2])).contiguous().view(-1)
    qc = M.QuadratureFunctionCoefficient(k)
    a = M.BilinearForm(S.fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(qc))
    a.AddDomainIntegrator(M.MassIntegrator(qc))
    a.Assemble()
    dd, dm = a.PAData()
    dd0, dm0 = S.a.PAData()
    assert _rel(S.host(dd), S.host(dd0)) < 1e-14
    assert _rel(S.host(dm), S.host(dm0)) < 1e-14
    assert _rel(S.host(dd), S.P.Dd) < RTOL
    x = np.random.default_rng(3 + p).uniform(-1, 1, S.P.ndofs)
    y = torch.empty(S.P.ndofs, dtype=torch.float64, device="cuda")
    a.Mult(S.dev(x), y)
    assert _rel(S.host(y), S.P.mult(x)) < RTOL
    bad = M.BilinearForm(S.fes)
    bad.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    bad.AddDomainIntegrator(M.DiffusionIntegrator(M.QuadratureFunctionCoefficient(k[:-1])))
    with pytest.raises(ValueError, match="QuadratureFunctionCoefficient"):
        bad.Assemble()


@pytest.mark.parametrize("p", range(1, 9))
def test_column_apply(orc, h1pa, torch_cuda, monkeypatch, p):
    """The fused column apply (opt-in with H1PA_COLUMNS=1: Mult of a diffusion form on a Cartesian mesh)
    on a mesh whose sizes are no multiple of the column cross-section, so
    columns are clipped in x and y and swept in several segments: within
    1e-13 of the oracle, within 1e-14 of the element-kernel + scatter path
    (H1PA_COLUMNS=0), the same bits on a second call, and AddMult adding the
    same result."""
    torch = torch_cuda
    monkeypatch.setenv("H1PA_COLUMNS", "1")  # opt-in path
    dims = (5, 7, 6) if p <= 4 else (3, 4, 5)
    S = Pair(orc, h1pa, torch, dims, p, perturb=0.2, seed=p, coeff="sincos")
    assert S.a.MultPath() == "columns"
    x = np.random.default_rng(20 + p).uniform(-1, 1, S.P.ndofs)
    xd = S.dev(x)
    y1 = torch.empty_like(xd)
    y2 = torch.empty_like(xd)
    S.a.Mult(xd, y1)
    S.a.Mult(xd, y2)
    yh = S.host(y1)
    assert _rel(yh, S.P.mult(x)) < RTOL
    assert np.array_equal(yh, S.host(y2)), "column apply is not deterministic"
    z = S.dev(np.ones(S.P.ndofs))
    S.a.AddMult(xd, z, 2.0)
    assert _rel(S.host(z), 1.0 + 2.0 * yh) < 1e-15
    monkeypatch.setenv("H1PA_COLUMNS", "0")
    a0 = h1pa.BilinearForm(S.fes)
    a0.SetAssemblyLevel(h1pa.AssemblyLevel.PARTIAL)
    a0.AddDomainIntegrator(h1pa.DiffusionIntegrator(h1pa.SinCosCoefficient()))
    a0.Assemble()
    y0 = torch.empty_like(xd)
    a0.Mult(xd, y0)
    assert _rel(yh, S.host(y0)) < 1e-14
