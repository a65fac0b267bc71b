"""GPU: edge cases of the operator and solver path against the oracle.

- Zero right-hand side: CGSolver's initial norm is 0, so it stops at
  iteration 0 with x = 0 (linalg/solvers.cpp:901-915); the same for the
  mixed-precision solver and for a zero solve inside a batch.
- No essential DOFs on an SPD operator (diffusion + mass): Jacobi-PCG
  iterations within +-1 of the oracle's, solution within 1e-9.
- The max_iter stop: not converged, MFEM's iteration count, and the oracle's
  iterate within 1e-12 after the same number of steps.
- One element, every order and both quadrature rules: apply within 1e-13.
"""
import numpy as np
import pytest

from test_gpu_parity import Pair

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _cg(M, S, A, ess, rel_tol=1e-12, max_iter=2000):
    cg = M.CGSolver()
    cg.SetRelTol(rel_tol)
    cg.SetMaxIter(max_iter)
    cg.SetOperator(A)
    cg.SetPreconditioner(M.OperatorJacobiSmoother(S.a, ess))
    return cg


@pytest.mark.parametrize("p", [2, 5])
def test_zero_rhs_stops_at_iteration_zero(orc, h1pa, torch_cuda, p):
    torch, M = torch_cuda, h1pa
    S = Pair(orc, M, torch, (3, 3, 2), p, perturb=0.1, seed=1, coeff="sincos")
    ess = S.fes.GetBoundaryTrueDofs()
    zero = torch.zeros(S.P.ndofs, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, zero, zero.clone())
    X.fill_(1.0)  # overwritten: iterative_mode is off
    cg = _cg(M, S, A, ess)
    cg.Mult(B, X)
    _, st = S.P.pcg(np.zeros(S.P.ndofs))
    assert cg.GetConverged() and st["converged"]
    assert cg.GetNumIterations() == st["iterations"] == 0
    assert not S.host(X).any()


def test_no_essential_dofs_mass_plus_diffusion(orc, h1pa, torch_cuda):
    torch, M = torch_cuda, h1pa
    S = Pair(orc, M, torch, (3, 2, 3), 3, perturb=0.15, seed=2, coeff="sincos",
             integrators=("diffusion", "mass"))
    ess = torch.zeros(0, dtype=torch.int32, device="cuda")
    b = np.random.default_rng(3).uniform(-1, 1, S.P.ndofs)
    A, X, B = S.a.FormLinearSystem(ess, torch.zeros(S.P.ndofs, dtype=torch.float64,
                                                    device="cuda"), S.dev(b))
    assert np.array_equal(S.host(B), b)
    cg = _cg(M, S, A, ess)
    cg.Mult(B, X)
    S.P.set_ess(np.zeros(0, dtype=np.int32))
    xo, st = S.P.pcg(b)
    print(f"[edge] no-ess iters gpu={cg.GetNumIterations()} oracle={st['iterations']}",
          flush=True)
    assert cg.GetConverged() and st["converged"]
    assert abs(cg.GetNumIterations() - st["iterations"]) <= 1
    assert _rel(S.host(X), xo) < 1e-9
    assert _rel(S.P.mult(S.host(X)), b) < 1e-9


@pytest.mark.parametrize("max_iter", [1, 5])
def test_max_iter_stop(orc, h1pa, torch_cuda, max_iter):
    torch, M = torch_cuda, h1pa
    S = Pair(orc, M, torch, (3, 3, 3), 3, perturb=0.2, seed=4, coeff="sincos")
    ess = S.fes.GetBoundaryTrueDofs()
    b = np.random.default_rng(5).uniform(-1, 1, S.P.ndofs)
    A, X, B = S.a.FormLinearSystem(ess, torch.zeros(S.P.ndofs, dtype=torch.float64,
                                                    device="cuda"), S.dev(b))
    cg = _cg(M, S, A, ess, max_iter=max_iter)
    cg.Mult(B, X)
    _, Bo = S.P.form_linear_system(np.zeros(S.P.ndofs), b)
    xo, st = S.P.pcg(Bo, max_iter=max_iter)
    assert not cg.GetConverged() and not st["converged"]
    assert cg.GetNumIterations() == st["iterations"]
    assert _rel(S.host(X), xo) < 1e-12


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("dq", [1, 2])
def test_single_element_apply(orc, h1pa, torch_cuda, p, dq):
    S = Pair(orc, h1pa, torch_cuda, (1, 1, 1), p, q1d=p + dq, coeff="sincos",
             integrators=("diffusion", "mass"))
    x = np.random.default_rng(p).uniform(-1, 1, S.P.ndofs)
    y = torch_cuda.empty(S.P.ndofs, dtype=torch_cuda.float64, device="cuda")
    S.a.Mult(S.dev(x), y)
    assert _rel(S.host(y), S.P.mult(x)) < 1e-13


def test_zero_rhs_in_mixed_and_batched_solves(orc, h1pa, torch_cuda):
    torch, M = torch_cuda, h1pa
    S = Pair(orc, M, torch, (3, 3, 3), 2, perturb=0.1, seed=6, coeff="sincos")
    ess = S.fes.GetBoundaryTrueDofs()
    zero = torch.zeros(S.P.ndofs, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, zero, zero.clone())
    ir = M.IRSolver(M.PREC_BF16_PA)
    ir.SetRelTol(1e-10)
    ir.SetMaxIter(200, 20000)
    ir.SetOperator(A)
    ir.SetPreconditioner(M.OperatorJacobiSmoother(S.a, ess))
    X.fill_(1.0)
    ir.Mult(B, X)
    assert ir.GetConverged()
    assert not S.host(X).any()
    # a zero solve next to a normal one in a mixed batch
    b = S.dev(np.random.default_rng(7).uniform(-1, 1, S.P.ndofs))
    A2, X2, B2 = S.a.FormLinearSystem(ess, zero, b)
    bs = M.BatchSolver(M.PREC_BF16_PA, iters_per_replay=8)
    X1 = torch.zeros_like(zero)
    bs.Add(A, B, X1, rel_tol=1e-10, max_iter=20000, tag=1)
    bs.Add(A2, B2, X2, rel_tol=1e-10, max_iter=20000, tag=2)
    assert bs.Run() == 0
    res = dict(bs.Poll())
    assert res[1]["converged"] and res[2]["converged"]
    assert not S.host(X1).any()
    assert res[2]["iterations"] > 0
