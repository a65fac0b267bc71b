"""GPU: p-refinement transfer (TensorProductPRefinementTransferOperator,
fem/transfer.cpp:2619-2990) and p-multigrid (MultigridBase, fem/multigrid.cpp:
106-240, levels as examples/ex26.cpp:42-110) through the C ABI, against the
oracle on the same mesh.

Bars: transfers within 1e-13 relative; one cycle with a fixed number of
coarse CG iterations (coarse rel_tol 0, so both sides run exactly that many)
within 1e-10; CG preconditioned by the cycle with MFEM's ex26 settings:
iteration count within +-1 of the oracle's and the solution within 1e-8.
"""
import numpy as np
import pytest

from test_gpu_mixed import _log_uniform, _rel

pytestmark = pytest.mark.gpu


def _mesh(M, dims, perturb, seed, curve, pg):
    mesh = M.Mesh.MakeCartesian3D(*dims)
    if perturb:
        mesh.Perturb(perturb, seed)
    if curve:
        mesh.SetCurvature(pg, curve)
    return mesh


@pytest.mark.parametrize("pl,ph", [(1, 2), (2, 4), (3, 5), (4, 8), (1, 3), (6, 7)])
def test_transfer_matches_oracle(orc, h1pa, torch_cuda, pl, ph):
    M, torch = h1pa, torch_cuda
    dims = (3, 4, 2)
    print(f"[case] transfer {pl}->{ph} dims={dims}", flush=True)
    mesh = _mesh(M, dims, 0.2, 9, 0.0, 1)
    lo = M.FiniteElementSpace(mesh, M.H1_FECollection(pl, 3))
    hi = M.FiniteElementSpace(mesh, M.H1_FECollection(ph, 3))
    Pl = orc.Problem(None, pl, perturb=0.2, seed=9, dims=dims)
    Ph = orc.Problem(None, ph, perturb=0.2, seed=9, dims=dims)
    To = orc.Transfer(Pl, Ph)
    T = M.TransferOperator(lo, hi)
    rng = np.random.default_rng(pl * 10 + ph)
    xl = rng.uniform(-1, 1, Pl.ndofs)
    xh = rng.uniform(-1, 1, Ph.ndofs)
    yh = torch.full((Ph.ndofs,), float("nan"), dtype=torch.float64, device="cuda")
    yl = torch.full((Pl.ndofs,), float("nan"), dtype=torch.float64, device="cuda")
    T.Mult(torch.from_numpy(xl).cuda(), yh)
    T.MultTranspose(torch.from_numpy(xh).cuda(), yl)
    torch.cuda.synchronize()
    assert _rel(yh.cpu().numpy(), To.mult(xl)) < 1e-13
    assert _rel(yl.cpu().numpy(), To.mult_transpose(xh)) < 1e-13


class MgCase:
    """A fine PA diffusion form (per-element log-uniform k, curved order-p
    geometry) and p-multigrid levels, built by the product and the oracle."""

    def __init__(self, orc, M, torch, dims, p, orders, seed, **mg):
        print(f"[case] multigrid dims={dims} p={p} orders={orders} {mg}", flush=True)
        ne = dims[0] * dims[1] * dims[2]
        k = _log_uniform(ne, seed)
        self.torch = torch
        self.mesh = _mesh(M, dims, 0.0, 0, 0.05, p)
        self.fes = M.FiniteElementSpace(self.mesh, M.H1_FECollection(p, 3))
        self.kd = torch.from_numpy(k).cuda()
        self.a = M.BilinearForm(self.fes)
        self.a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
        self.a.AddDomainIntegrator(M.DiffusionIntegrator(M.PWConstCoefficient(self.kd)))
        self.a.Assemble()
        self.mg = M.PMultigrid(self.a, orders, **mg)
        fine = orc.Problem(None, p, curve=0.05, coeff=k, dims=dims)
        levels = [orc.Problem(None, q, coeff=k, dims=dims, nodes=(p, fine.X))
                  for q in orders[:-1]] + [fine]
        omg = {"smoother_order": mg.get("smoother_order", 2),
               "coarse_jacobi": mg.get("coarse_jacobi", False),
               "coarse_rel_tol": mg.get("coarse_rel_tol", 1e-2),
               "coarse_max_iter": mg.get("coarse_max_iter", 200),
               "pre": mg.get("pre_smoothing", 1), "post": mg.get("post_smoothing", 1),
               "w_cycle": mg.get("cycle", "V").upper().startswith("W")}
        self.omg = orc.PMultigrid(levels, **omg)
        self.P = fine
        assert self.fes.GetNDofs() == fine.ndofs


def test_cycle_matches_oracle(orc, h1pa, torch_cuda):
    torch = torch_cuda
    S = MgCase(orc, h1pa, torch, (3, 3, 4), 4, (1, 2, 4), 31, coarse_rel_tol=0.0,
               coarse_max_iter=6)
    info = S.mg.Info()
    assert np.allclose(info["max_eig"][1:], S.omg.max_eig[1:], rtol=1e-10)
    x = np.random.default_rng(2).uniform(-1, 1, S.P.ndofs)
    y = torch.empty(S.P.ndofs, dtype=torch.float64, device="cuda")
    S.mg.Mult(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy(), S.omg.mult(x)) < 1e-10
    info = S.mg.Info()
    assert info["cycles"] == 1 and info["coarse_solves"] == 1 and info["coarse_iterations"] == 6


def test_w_cycle_matches_oracle(orc, h1pa, torch_cuda):
    """W-cycle, two pre- and one post-smoothing step, order-3 Chebyshev, four
    levels; fixed coarse iterations."""
    torch = torch_cuda
    S = MgCase(orc, h1pa, torch, (3, 2, 3), 6, (1, 2, 3, 6), 17, coarse_rel_tol=0.0,
               coarse_max_iter=5, cycle="W", pre_smoothing=2, post_smoothing=1,
               smoother_order=3)
    x = np.random.default_rng(8).uniform(-1, 1, S.P.ndofs)
    y = torch.empty(S.P.ndofs, dtype=torch.float64, device="cuda")
    S.mg.Mult(torch.from_numpy(x).cuda(), y)
    torch.cuda.synchronize()
    assert _rel(y.cpu().numpy(), S.omg.mult(x)) < 1e-10
    info = S.mg.Info()
    # a W-cycle visits level l 2^(top - l) times: the coarsest 2^3 = 8 times
    assert info["coarse_solves"] == 8 and info["coarse_iterations"] == 40


@pytest.mark.parametrize("p,orders,cj", [(4, (1, 2, 4), False), (6, (1, 3, 6), False),
                                         (3, (1, 3), False), (4, (1, 2, 4), True)])
def test_multigrid_pcg_matches_oracle(orc, h1pa, torch_cuda, p, orders, cj):
    torch, M = torch_cuda, h1pa
    dims = (4, 4, 4) if p <= 4 else (3, 3, 3)
    S = MgCase(orc, M, torch, dims, p, orders, 40 + p, coarse_jacobi=cj)
    b = np.random.default_rng(p).uniform(-1, 1, S.P.ndofs)
    ess = S.fes.GetBoundaryTrueDofs()
    x0 = torch.zeros(S.P.ndofs, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, x0, torch.from_numpy(b).cuda())
    Xo, Bo = S.P.form_linear_system(np.zeros(S.P.ndofs), b)
    cg = M.CGSolver()
    cg.SetRelTol(1e-10)
    cg.SetMaxIter(500)
    cg.SetOperator(A)
    cg.SetPreconditioner(S.mg)
    cg.Mult(B, X)
    xo, it_o, conv_o = S.omg.pcg(Bo, rel_tol=1e-10, max_iter=500)
    it = cg.GetNumIterations()
    x = X.cpu().numpy()
    print(f"[mg] p={p} orders={orders} iterations {it} (oracle {it_o}), "
          f"rel diff {_rel(x, xo):.3e}", flush=True)
    assert cg.GetConverged() and conv_o
    assert abs(it - it_o) <= 1
    assert _rel(x, xo) < 1e-8
    # and fewer iterations than Jacobi-PCG on the same system (the gap grows
    # with the mesh: these meshes have 3-4 elements per direction)
    _, st = S.P.pcg(Bo, rel_tol=1e-10, max_iter=20000)
    assert it < st["iterations"]


def test_multigrid_argument_errors(h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    mesh = M.Mesh.MakeCartesian3D(2, 2, 2)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(3, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.ConstantCoefficient(1.0)))
    a.Assemble()
    for orders in ((1, 2), (2, 1, 3), (3,), (0, 3)):
        with pytest.raises(M.H1paError):
            M.PMultigrid(a, orders)
    k = torch.ones(8 * 125, dtype=torch.float64, device="cuda")
    aq = M.BilinearForm(fes)
    aq.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    aq.AddDomainIntegrator(M.DiffusionIntegrator(M.QuadratureFunctionCoefficient(k)))
    aq.Assemble()
    with pytest.raises(M.H1paError) as e:
        M.PMultigrid(aq, (1, 3))
    assert e.value.code == M.H1PA_EUNSUP
    # the hierarchy holds the fine form: re-assembly and destroy are refused
    # until the hierarchy is destroyed
    mg = M.PMultigrid(a, (1, 3))
    with pytest.raises(M.H1paError) as e:
        a.Assemble()
    assert e.value.code == M.H1PA_EINVAL and "multigrid" in str(e.value)
    assert M.lib().h1pa_form_destroy(a._single()._h) == M.H1PA_EINVAL
    del mg
    a.Assemble()


def test_bf16_cycle_matches_oracle(orc, h1pa, torch_cuda):
    """A mixed-precision cycle (every level's applies read BF16 copies of its
    PA data; FP64 vectors, diagonals from the FP64 data) inside the FP64 CG:
    the oracle runs the same cycle on BF16-rounded PA data; iterations within
    +-1, solution within 1e-8, and the FP64 true residual meets the stop."""
    torch, M = torch_cuda, h1pa
    dims, p, orders = (4, 4, 4), 4, (1, 2, 4)
    ne = 64
    k = _log_uniform(ne, 77)
    mesh = _mesh(M, dims, 0.0, 0, 0.05, p)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    kd = torch.from_numpy(k).cuda()
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.PWConstCoefficient(kd)))
    a.Assemble()
    mg = M.PMultigrid(a, orders, cycle_precision=M.PREC_BF16_PA)
    fine = orc.Problem(None, p, curve=0.05, coeff=k, dims=dims)
    lv = [orc.Problem(None, q, coeff=k, dims=dims, nodes=(p, fine.X)) for q in orders]
    diags = [P.diagonal() for P in lv]
    for P in lv:
        P.replace_pa(Dd=orc.round_bf16(P.Dd))
    omg = orc.PMultigrid(lv, diags=diags)
    b = np.random.default_rng(3).uniform(-1, 1, fine.ndofs)
    ess = fes.GetBoundaryTrueDofs()
    A, X, B = a.FormLinearSystem(ess, torch.zeros(fine.ndofs, dtype=torch.float64,
                                                  device="cuda"), torch.from_numpy(b).cuda())
    _, Bo = fine.form_linear_system(np.zeros(fine.ndofs), b)
    cg = M.CGSolver()
    cg.SetRelTol(1e-10)
    cg.SetMaxIter(500)
    cg.SetOperator(A)
    cg.SetPreconditioner(mg)
    cg.Mult(B, X)
    xo, it_o, conv_o = orc.cg_preconditioned(fine, omg.mult, Bo, 1e-10, 0.0, 500)
    it = cg.GetNumIterations()
    x = X.cpu().numpy()
    print(f"[mg-bf16] iterations {it} (oracle {it_o}), rel diff {_rel(x, xo):.3e}", flush=True)
    assert cg.GetConverged() and conv_o
    assert abs(it - it_o) <= 1
    assert _rel(x, xo) < 1e-8
    r = Bo - fine.constrained_mult(x)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(Bo)
