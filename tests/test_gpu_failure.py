"""GPU: CGSolver's failure outcomes (linalg/solvers.cpp:869-1050) in the
single, mixed-precision and batched solves.

  - A non-finite (B r, r), (A d, d) or betanom makes MFEM_VERIFY(IsFinite(...))
    abort CGSolver::Mult; here the solve stops with H1PA_ENONFINITE (an error,
    as mfem_error is), and in a batch that solve alone retires with that
    status while the others run on, bitwise as they run alone.
  - A preconditioner that is not positive definite ((B r, r) < 0) ends the
    solve at once, not converged, as CGSolver does; the oracle's PCG agrees on
    the outcome and the iteration count.
These are numeric conditions set up on the host (a NaN in the right-hand
side, a negative coefficient); every kernel runs to its normal end.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(M, torch, n, p, seed, k=None):
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.15, seed)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    coef = M.FourierCoefficient(seed) if k is None else M.PWConstCoefficient(k)
    a.AddDomainIntegrator(M.DiffusionIntegrator(coef))
    a.Assemble()
    N = fes.GetNDofs()
    b = torch.from_numpy(np.random.default_rng(seed).uniform(-1, 1, N)).cuda()
    ess = fes.GetBoundaryTrueDofs()
    A, X, B = a.FormLinearSystem(ess, torch.zeros_like(b), b)
    return dict(mesh=mesh, fes=fes, a=a, A=A, X=X, B=B, ess=ess, N=N, seed=seed)


def _free_dof(P):
    ess = set(P["ess"].cpu().numpy().tolist())
    return next(i for i in range(P["N"]) if i not in ess)


def test_nan_rhs_single_and_mixed(h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    P = _problem(M, torch, 4, 3, 1)
    P["B"][_free_dof(P)] = float("nan")
    cg = M.CGSolver()
    cg.SetRelTol(1e-10)
    cg.SetMaxIter(500)
    cg.SetOperator(P["A"])
    cg.SetPreconditioner(M.OperatorJacobiSmoother(P["a"], P["ess"]))
    with pytest.raises(M.H1paError) as ei:
        cg.Mult(P["B"], P["X"].clone())
    assert ei.value.code == M.H1PA_ENONFINITE
    assert not cg.GetConverged()
    ir = M.IRSolver(M.PREC_FP32)
    ir.SetRelTol(1e-10)
    ir.SetOperator(P["A"])
    with pytest.raises(M.H1paError) as ei:
        ir.Mult(P["B"], P["X"].clone())
    assert ei.value.code == M.H1PA_ENONFINITE


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_nan_slot_retires_alone(h1pa, torch_cuda, prec):
    M, torch = h1pa, torch_cuda
    pr = M.PREC_FP64 if prec == "fp64" else M.PREC_BF16_PA
    probs = [_problem(M, torch, n, p, s) for (n, p, s) in [(4, 2, 11), (3, 4, 12), (5, 3, 13)]]
    bad = probs[1]
    bad["B"][_free_dof(bad)] = float("nan")
    alone = {}
    for P in (probs[0], probs[2]):
        bs = M.BatchSolver(pr)
        x = P["X"].clone()
        bs.Add(P["A"], P["B"], x, rel_tol=1e-10, max_iter=5000, tag=P["seed"])
        assert bs.Run() == 0
        alone[P["seed"]] = (x.cpu().numpy(), dict(bs.Poll())[P["seed"]])
    bs = M.BatchSolver(pr)
    xs = {}
    for P in probs:
        xs[P["seed"]] = P["X"].clone()
        bs.Add(P["A"], P["B"], xs[P["seed"]], rel_tol=1e-10, max_iter=5000, tag=P["seed"])
    assert bs.Run() == 0, "a failed solve must not keep the batch running"
    res = dict(bs.Poll())
    print(res, flush=True)
    assert res[12]["status"] == M.H1PA_ENONFINITE and not res[12]["converged"]
    for s in (11, 13):
        assert res[s]["status"] == M.H1PA_OK and res[s]["converged"]
        assert res[s] == alone[s][1]
        assert np.array_equal(xs[s].cpu().numpy(), alone[s][0])


def test_indefinite_preconditioner_stops(orc, h1pa, torch_cuda):
    """k = -1 on every element: the Jacobi diagonal is negative, (B r, r) < 0
    at the start, and CGSolver returns at once, not converged (final_iter 0),
    in the single solve, the batch and the oracle."""
    M, torch = h1pa, torch_cuda
    n, p = 3, 2
    k = -np.ones(n ** 3)
    P = _problem(M, torch, n, p, 21, k=torch.from_numpy(k).cuda())
    cg = M.CGSolver()
    cg.SetRelTol(1e-10)
    cg.SetMaxIter(500)
    cg.SetOperator(P["A"])
    cg.SetPreconditioner(M.OperatorJacobiSmoother(P["a"], P["ess"]))
    cg.Mult(P["B"], P["X"].clone())
    assert not cg.GetConverged() and cg.GetNumIterations() == 0
    bs = M.BatchSolver(M.PREC_FP64)
    bs.Add(P["A"], P["B"], P["X"].clone(), rel_tol=1e-10, max_iter=500, tag=1)
    assert bs.Run() == 0
    r = dict(bs.Poll())[1]
    assert r["status"] == M.H1PA_ENOTCONV and not r["converged"] and r["iterations"] == 0
    Po = orc.Problem(n, p, perturb=0.15, seed=21, coeff=k)
    b = P["B"].cpu().numpy()
    _, st = Po.pcg(b, rel_tol=1e-10, max_iter=500)
    assert not st["converged"] and st["iterations"] == 0
