"""GPU: OperatorChebyshevSmoother and Chebyshev-preconditioned CG against the
oracle (linalg/solvers.cpp:455-656, linalg/operator.cpp:886-935).

Bars, written per test:
  - eigenvalue estimate (10 power steps from Vector::Randomize(12345)):
    relative 1e-10 of the oracle's; the start vector must be the same draw
    for this to hold, since 10 steps leave the estimate far from converged;
  - coefficients from a given bound: relative 1e-14;
  - Mult: relative 1e-12 in the 2-norm;
  - PCG: iteration count within +-1 of the oracle's, solution within 1e-9
    (rel_tol 1e-12).
"""
import numpy as np
import pytest

from test_gpu_parity import Pair

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _system(S, torch, seed=5):
    b = np.random.default_rng(seed).uniform(-1, 1, S.P.ndofs)
    ess = S.fes.GetBoundaryTrueDofs()
    x0 = torch.zeros(S.P.ndofs, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, x0, S.dev(b))
    _, Bo = S.P.form_linear_system(np.zeros(S.P.ndofs), b)
    return ess, A, X, B, Bo


@pytest.mark.parametrize("p", [2, 3, 4])
def test_power_method_estimate_matches_oracle(orc, h1pa, torch_cuda, p):
    S = Pair(orc, h1pa, torch_cuda, (4, 3, 3), p, perturb=0.2, seed=p, coeff="sincos")
    ess, A, *_ = _system(S, torch_cuda)
    cheb = h1pa.OperatorChebyshevSmoother(A, None, ess, 3)
    eo, co = S.P.chebyshev(3)
    e = cheb.GetMaxEigEstimate()
    print(f"[cheb] p={p} eig gpu={e:.15g} oracle={eo:.15g}", flush=True)
    assert abs(e - eo) <= 1e-10 * abs(eo)
    assert np.allclose(cheb.GetCoefficients(), co, rtol=1e-9, atol=0)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_coefficients_and_mult_match_oracle(orc, h1pa, torch_cuda, order):
    S = Pair(orc, h1pa, torch_cuda, (3, 3, 3), 3, perturb=0.2, seed=7, coeff="sincos")
    ess, A, *_ = _system(S, torch_cuda)
    lam = 2.345678
    cheb = h1pa.OperatorChebyshevSmoother(A, None, ess, order, max_eig_estimate=lam)
    co = orc.chebyshev_coeffs(order, lam)
    assert cheb.GetMaxEigEstimate() == lam
    assert np.allclose(cheb.GetCoefficients(), co, rtol=1e-14, atol=0)
    x = np.random.default_rng(order).uniform(-1, 1, S.P.ndofs)
    y = torch_cuda.empty(S.P.ndofs, dtype=torch_cuda.float64, device="cuda")
    cheb.Mult(S.dev(x), y)
    yo = S.P.chebyshev_mult(co, x)
    err = _rel(S.host(y), yo)
    print(f"[cheb] order={order} mult rel err={err:.2e}", flush=True)
    assert err <= 1e-12


def test_caller_diagonal_and_own_ess_list(orc, h1pa, torch_cuda):
    """A caller-provided diagonal (MFEM's `d`) and an essential list that is
    not the space's own (copied by the smoother)."""
    torch = torch_cuda
    S = Pair(orc, h1pa, torch, (3, 3, 2), 2, perturb=0.1, seed=4, coeff="sincos")
    ess = S.fes.GetBoundaryTrueDofs()
    sub = ess[::2].clone()
    x0 = torch.zeros(S.P.ndofs, dtype=torch.float64, device="cuda")
    b = np.random.default_rng(2).uniform(-1, 1, S.P.ndofs)
    A, X, B = S.a.FormLinearSystem(sub, x0, S.dev(b))
    diag = torch.empty(S.P.ndofs, dtype=torch.float64, device="cuda")
    S.a.AssembleDiagonal(diag)
    c1 = h1pa.OperatorChebyshevSmoother(A, diag, sub, 2, max_eig_estimate=1.7)
    c2 = h1pa.OperatorChebyshevSmoother(A, None, sub, 2, max_eig_estimate=1.7)
    v = S.dev(np.random.default_rng(3).uniform(-1, 1, S.P.ndofs))
    y1 = torch.empty_like(v)
    y2 = torch.empty_like(v)
    c1.Mult(v, y1)
    sub.zero_()  # the smoother keeps its own copy of the list
    c2.Mult(v, y2)
    assert torch.equal(y1, y2)


@pytest.mark.parametrize("p,order", [(2, 2), (3, 3), (4, 2)])
def test_pcg_chebyshev_matches_oracle(orc, h1pa, torch_cuda, p, order):
    S = Pair(orc, h1pa, torch_cuda, (5, 4, 4), p, perturb=0.2, seed=11 + p, coeff="sincos")
    ess, A, X, B, Bo = _system(S, torch_cuda, seed=p)
    cheb = h1pa.OperatorChebyshevSmoother(A, None, ess, order)
    cg = h1pa.CGSolver()
    cg.SetRelTol(1e-12)
    cg.SetMaxIter(2000)
    cg.SetOperator(A)
    cg.SetPreconditioner(cheb)
    cg.Mult(B, X)
    eo, co = S.P.chebyshev(order)
    xo, st = S.P.pcg_chebyshev(co, Bo, rel_tol=1e-12, max_iter=2000)
    _, sj = S.P.pcg(Bo, rel_tol=1e-12, max_iter=2000)
    print(f"[chebpcg] p={p} order={order} gpu iters={cg.GetNumIterations()} oracle="
          f"{st['iterations']} jacobi={sj['iterations']}", flush=True)
    assert cg.GetConverged() and st["converged"]
    assert abs(cg.GetNumIterations() - st["iterations"]) <= 1
    assert _rel(S.host(X), xo) < 1e-9
    assert st["iterations"] < sj["iterations"]


def test_pcg_chebyshev_large_converges(h1pa, torch_cuda):
    """~2M-DOF p = 4 curved mesh: Chebyshev(3)-PCG reaches 1e-10 on the true
    residual measure in fewer iterations than Jacobi-PCG."""
    torch = torch_cuda
    M = h1pa
    n, p = 32, 4
    print(f"[case] chebyshev large n={n} p={p}", flush=True)
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.SetCurvature(p, 0.05)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient()))
    a.Assemble()
    N = fes.GetNDofs()
    b = torch.empty(N, dtype=torch.float64, device="cuda")
    M.Randomize(b, 77, -1.0, 1.0)
    ess = fes.GetBoundaryTrueDofs()
    A, X, B = a.FormLinearSystem(ess, torch.zeros(N, dtype=torch.float64, device="cuda"), b)
    its = {}
    for name in ("jacobi", "chebyshev"):
        cg = M.CGSolver()
        cg.SetRelTol(1e-10)
        cg.SetMaxIter(5000)
        cg.SetOperator(A)
        prec = (M.OperatorJacobiSmoother(a, ess) if name == "jacobi"
                else M.OperatorChebyshevSmoother(A, None, ess, 3))
        cg.SetPreconditioner(prec)
        Xs = X.clone()
        cg.Mult(B, Xs)
        assert cg.GetConverged()
        y = torch.empty_like(B)
        A.Mult(Xs, y)
        res = (torch.linalg.norm(B - y) / torch.linalg.norm(B)).item()
        its[name] = cg.GetNumIterations()
        print(f"[chebpcg] {name} iters={its[name]} rel res={res:.2e}", flush=True)
        assert res < 1e-7
    assert its["chebyshev"] < its["jacobi"]


def test_bad_arguments(orc, h1pa, torch_cuda):
    S = Pair(orc, h1pa, torch_cuda, (2, 2, 2), 2)
    ess, A, *_ = _system(S, torch_cuda)
    with pytest.raises(h1pa.H1paError):
        h1pa.OperatorChebyshevSmoother(A, None, ess, 6)
    with pytest.raises(h1pa.H1paError):
        h1pa.OperatorChebyshevSmoother(A, None, ess, 2, power_seed=0)
