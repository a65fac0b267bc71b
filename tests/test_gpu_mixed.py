"""GPU: reduced-precision applies and the mixed-precision iterative-refinement
solver (no MFEM counterpart; checked against the FP64 oracle).

Bars, written per test:
  - FP32 apply vs the FP64 oracle apply: 1e-5 relative (norm-wise).
  - BF16-stored PA apply vs the oracle applied with the same BF16-rounded PA
    data in FP64: 1e-5 relative; vs the unrounded oracle: the BF16 storage
    error (between 1e-5 and 3e-2), which shows the BF16 data is what runs.
  - Mixed solve: CGSolver's measure sqrt((B r, r) / (B b, b)) of the FP64
    true residual, recomputed by the oracle, <= 1.05e-10 (the stop test at
    1e-10, 5% for the different summation order), and the solution within
    1e-6 relative of the oracle's FP64 PCG solution.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _measure(P, x, b, constrained=True):
    """sqrt((B r, r) / (B b, b)), r = b - A x, B = Jacobi (DIAG_ONE rows 1)."""
    r = b - (P.constrained_mult(x) if constrained else P.mult(x))
    dinv = 1.0 / P.diagonal()
    if constrained:
        dinv[P.ess] = 1.0
    return np.sqrt(r @ (dinv * r)) / np.sqrt(b @ (dinv * b))


def _log_uniform(ne, seed, lo=1e-2, hi=1e2):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), ne))


class Case:
    def __init__(self, orc, M, torch, dims, p, perturb=0.0, seed=0, curve=0.0, coeff="one",
                 integrators=("diffusion",)):
        print(f"[case] dims={dims} p={p} perturb={perturb} curve={curve} "
              f"coeff={'array' if not isinstance(coeff, str) else coeff} integ={integrators}",
              flush=True)
        self.torch, self.M = torch, M
        self.mesh = M.Mesh.MakeCartesian3D(*dims)
        if perturb:
            self.mesh.Perturb(perturb, seed)
        if curve:
            self.mesh.SetCurvature(p, curve)
        self.fes = M.FiniteElementSpace(self.mesh, M.H1_FECollection(p, 3))
        self.a = M.BilinearForm(self.fes)
        self.a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
        if isinstance(coeff, str):
            coef = M.SinCosCoefficient() if coeff == "sincos" else M.ConstantCoefficient(1.0)
        else:
            self._cdev = torch.from_numpy(np.ascontiguousarray(coeff)).cuda()
            coef = M.PWConstCoefficient(self._cdev)
        for it in integrators:
            cls = M.DiffusionIntegrator if it == "diffusion" else M.MassIntegrator
            self.a.AddDomainIntegrator(cls(coef))
        self.a.Assemble()
        self.P = orc.Problem(None, p, perturb=perturb, seed=seed, curve=curve, coeff=coeff,
                             integrators=integrators, dims=dims)
        assert self.fes.GetNDofs() == self.P.ndofs
        self.n = self.P.ndofs

    def dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).cuda()

    def host(self, t):
        self.torch.cuda.synchronize()
        return t.cpu().numpy()


def _dims(p):
    return (3, 2, 3) if p <= 5 else (2, 2, 2)


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("integ", [("diffusion",), ("mass",), ("diffusion", "mass")])
def test_mult_fp32(orc, h1pa, torch_cuda, p, integ):
    S = Case(orc, h1pa, torch_cuda, _dims(p), p, perturb=0.2, seed=5, coeff="sincos",
             integrators=integ)
    x32 = np.random.default_rng(p).uniform(-1, 1, S.n).astype(np.float32)
    y = torch_cuda.empty(S.n, dtype=torch_cuda.float32, device="cuda")
    S.a.MultLowPrecision(S.dev(x32), y, h1pa.PREC_FP32)
    yo = S.P.mult(x32.astype(np.float64))
    err = _rel(S.host(y).astype(np.float64), yo)
    print(f"[fp32] p={p} {integ} rel={err:.3e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5, 6, 7, 8])
@pytest.mark.parametrize("integ", [("diffusion",), ("diffusion", "mass")])
def test_mult_bf16_pa(orc, h1pa, torch_cuda, p, integ):
    S = Case(orc, h1pa, torch_cuda, _dims(p), p, perturb=0.2, seed=6, coeff="sincos",
             integrators=integ)
    x32 = np.random.default_rng(10 + p).uniform(-1, 1, S.n).astype(np.float32)
    x64 = x32.astype(np.float64)
    y = torch_cuda.empty(S.n, dtype=torch_cuda.float32, device="cuda")
    S.a.MultLowPrecision(S.dev(x32), y, h1pa.PREC_BF16_PA)
    yg = S.host(y).astype(np.float64)
    yo = S.P.mult(x64)
    S.P.replace_pa(Dd=orc.round_bf16(S.P.Dd) if S.P.Dd is not None else None,
                   Mv=orc.round_bf16(S.P.Mv) if S.P.Mv is not None else None)
    yr = S.P.mult(x64)
    e_same, e_full = _rel(yg, yr), _rel(yg, yo)
    print(f"[bf16] p={p} {integ} vs rounded-PA oracle {e_same:.3e}, vs FP64 {e_full:.3e}",
          flush=True)
    assert e_same < 1e-5
    assert 1e-5 < e_full < 3e-2


def test_fp32_then_bf16_then_fp32(orc, h1pa, torch_cuda):
    """Switching the shadow precision rebuilds it: results do not leak across."""
    S = Case(orc, h1pa, torch_cuda, (2, 3, 2), 4, perturb=0.1, seed=2, coeff="sincos")
    x32 = np.random.default_rng(0).uniform(-1, 1, S.n).astype(np.float32)
    y1, y2, y3 = (torch_cuda.empty(S.n, dtype=torch_cuda.float32, device="cuda")
                  for _ in range(3))
    S.a.MultLowPrecision(S.dev(x32), y1, h1pa.PREC_FP32)
    S.a.MultLowPrecision(S.dev(x32), y2, h1pa.PREC_BF16_PA)
    S.a.MultLowPrecision(S.dev(x32), y3, h1pa.PREC_FP32)
    a1, a2, a3 = S.host(y1), S.host(y2), S.host(y3)
    assert np.array_equal(a1, a3)
    assert not np.array_equal(a1, a2)


def _solve_case(orc, h1pa, torch, dims, p, curve, seed, integ=("diffusion",)):
    ne = dims[0] * dims[1] * dims[2]
    S = Case(orc, h1pa, torch, dims, p, curve=curve, coeff=_log_uniform(ne, seed),
             integrators=integ)
    b = np.random.default_rng(seed + 1).uniform(-1, 1, S.n)
    ess = S.fes.GetBoundaryTrueDofs()
    x0 = torch.zeros(S.n, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, x0, S.dev(b))
    Xo, Bo = S.P.form_linear_system(np.zeros(S.n), b)
    assert _rel(S.host(B), Bo) < 1e-13
    xo, st = S.P.pcg(Bo, rel_tol=1e-13, max_iter=20000)
    assert st["converged"]
    return S, ess, A, X, B, Bo, xo


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("p", [2, 4, 6])
def test_ir_solve_curved_jumps(orc, h1pa, torch_cuda, p, prec):
    """cfg3 shape at small size: curved map (amplitude 0.05), geometry of
    order p, per-element log-uniform coefficient in [1e-2, 1e2], RHS
    uniform[-1,1], homogeneous Dirichlet, IR to 1e-10."""
    torch = torch_cuda
    dims = (5, 5, 5) if p <= 4 else (4, 4, 4)
    S, ess, A, X, B, Bo, xo = _solve_case(orc, h1pa, torch, dims, p, 0.05, 40 + p)
    st_o = S.P.pcg(Bo, rel_tol=1e-10, max_iter=20000)[1]["iterations"]
    P = h1pa.PREC_FP32 if prec == "fp32" else h1pa.PREC_BF16_PA
    ir = h1pa.IRSolver(P)
    ir.SetRelTol(1e-10)
    ir.SetMaxIter(200, 20000)
    ir.SetOperator(A)
    ir.SetPreconditioner(h1pa.OperatorJacobiSmoother(S.a, ess))
    ir.Mult(B, X)
    st = ir.stats
    x = S.host(X)
    meas = _measure(S.P, x, Bo)
    err = _rel(x, xo)
    print(f"[ir] p={p} {prec} outer={st.outer_iterations} inner={st.inner_iterations} "
          f"oracle_pcg_iters={st_o} meas={meas:.3e} res2={st.rel_residual:.3e} "
          f"err_vs_fp64={err:.3e}", flush=True)
    assert ir.GetConverged()
    assert meas <= 1.05e-10
    assert err <= 1e-6


def test_ir_solve_mass_plus_diffusion_unconstrained(orc, h1pa, torch_cuda):
    """No essential DOFs (mass + diffusion is SPD alone)."""
    torch = torch_cuda
    dims, p = (3, 4, 3), 3
    ne = 36
    S = Case(orc, h1pa, torch, dims, p, perturb=0.15, seed=8, coeff=_log_uniform(ne, 8),
             integrators=("diffusion", "mass"))
    b = np.random.default_rng(9).uniform(-1, 1, S.n)
    ess = torch.zeros(0, dtype=torch.int32, device="cuda")
    x0 = torch.zeros(S.n, dtype=torch.float64, device="cuda")
    A, X, B = S.a.FormLinearSystem(ess, x0, S.dev(b))
    ir = h1pa.IRSolver(h1pa.PREC_BF16_PA)
    ir.SetRelTol(1e-10)
    ir.SetOperator(A)
    ir.Mult(B, X)
    x = S.host(X)
    meas = _measure(S.P, x, b, constrained=False)
    print(f"[ir] mass+diffusion outer={ir.stats.outer_iterations} meas={meas:.3e}", flush=True)
    assert ir.GetConverged() and meas <= 1.05e-10


def test_ir_not_converged_reports(orc, h1pa, torch_cuda):
    """One refinement step cannot reach 1e-14: the call reports no convergence
    (H1PA_ENOTCONV, mapped to GetConverged() False) and a finite residual."""
    torch = torch_cuda
    S, ess, A, X, B, Bo, xo = _solve_case(orc, h1pa, torch, (3, 3, 3), 2, 0.0, 3)
    ir = h1pa.IRSolver(h1pa.PREC_FP32)
    ir.SetRelTol(1e-14)
    ir.SetDelta(0.5)
    ir.SetMaxIter(1, 50)
    ir.SetOperator(A)
    ir.Mult(B, X)
    assert not ir.GetConverged()
    assert ir.stats.outer_iterations == 1
    assert 0 < ir.stats.rel_residual < 1


def test_ir_matches_fp64_pcg(orc, h1pa, torch_cuda):
    """The mixed solve and the FP64 GPU PCG (rel_tol 1e-10) agree with the
    oracle solution to what the residual bound allows."""
    torch = torch_cuda
    S, ess, A, X, B, Bo, xo = _solve_case(orc, h1pa, torch, (4, 4, 4), 4, 0.05, 77)
    cg = h1pa.CGSolver()
    cg.SetRelTol(1e-10)
    cg.SetMaxIter(5000)
    cg.SetOperator(A)
    cg.SetPreconditioner(h1pa.OperatorJacobiSmoother(S.a, ess))
    X64 = X.clone()
    cg.Mult(B, X64)
    assert cg.GetConverged()
    ir = h1pa.IRSolver(h1pa.PREC_BF16_PA)
    ir.SetRelTol(1e-10)
    ir.SetOperator(A)
    ir.Mult(B, X)
    x64, xir = S.host(X64), S.host(X)
    print(f"[ir] pcg iters={cg.GetNumIterations()} ir inner={ir.stats.inner_iterations} "
          f"err pcg={_rel(x64, xo):.3e} ir={_rel(xir, xo):.3e}", flush=True)
    assert ir.GetConverged()
    assert _rel(xir, xo) <= 1e-6 and _rel(x64, xo) <= 1e-6
