"""GPU parity at the configurations' own sizes (the HIP path through the C
ABI against the CPU oracle).

  - Operator-apply sweep at about 1M DOFs for every order p = 1..8 (n = 99,
    49, 33, 25, 20, 16, 14, 12: mostly not powers of two, so the
    space-filling-curve order and the element-block tails are the general
    ones), interior vertices perturbed by 0.2h, the sin/cos coefficient at the
    physical quadrature points, the apply-variant autotuner on: restriction
    maps bit-exact; PA diffusion and PA mass Mult and AssembleDiagonal within
    1e-13 relative.  Reference: fem/restriction.cpp:26-228,
    fem/integ/bilininteg_diffusion_kernels.hpp:989-1221,
    fem/integ/bilininteg_mass_kernels.hpp:815-1049.
  - The benchmark's own workload (79^3, p = 4, 31.9M DOFs): Mult within 1e-13.
  - The batched workload's problems (n ~ U[4, 24], p ~ U{2..6}, 0.15h
    perturbation, k = exp(8 seeded Fourier modes)): PA data of the Fourier
    coefficient within 1e-13, and FP64 and BF16-mixed batches against the
    oracle's FP64 Jacobi-PCG (iterations within +-1, FP64 solutions within
    1e-8; mixed: CGSolver's measure of the FP64 true residual <= 1.05e-10).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-13
N_1M = {1: 99, 2: 49, 3: 33, 4: 25, 5: 20, 6: 16, 7: 14, 8: 12}


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def _form(M, fes, integ, coeff):
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator((M.DiffusionIntegrator if integ == "diffusion" else M.MassIntegrator)(coeff))
    a.Assemble()
    return a


@pytest.mark.parametrize("p", range(1, 9))
def test_cfg2_1M_maps_and_operators(orc, h1pa, torch_cuda, p):
    M, torch = h1pa, torch_cuda
    n = N_1M[p]
    print(f"[scale] cfg2 p={p} n={n}^3", flush=True)
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, p)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    Pd = orc.Problem(n, p, perturb=0.2, seed=p, coeff="sincos", integrators=("diffusion",))
    assert fes.GetNDofs() == Pd.ndofs
    R = fes.GetElementRestriction()
    assert np.array_equal(R.GatherMap().cpu().numpy(), Pd.gather_map)
    assert np.array_equal(R.Offsets().cpu().numpy(), Pd.offsets)
    assert np.array_equal(R.Indices().cpu().numpy(), Pd.indices)
    assert np.array_equal(fes.GetBoundaryTrueDofs().cpu().numpy(), Pd.ess)
    x = np.random.default_rng(p).uniform(-1, 1, Pd.ndofs)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    for integ in ("diffusion", "mass"):
        P = Pd if integ == "diffusion" else orc.Problem(n, p, perturb=0.2, seed=p, coeff="sincos",
                                                        integrators=("mass",))
        a = _form(M, fes, integ, M.SinCosCoefficient())
        a.Mult(xd, y)
        torch.cuda.synchronize()
        e1 = _rel(y.cpu().numpy(), P.mult(x))
        a.AssembleDiagonal(y)
        torch.cuda.synchronize()
        e2 = _rel(y.cpu().numpy(), P.diagonal())
        print(f"[scale] {integ} variant={a.Info()['variant']} mult {e1:.2e} diag {e2:.2e}",
              flush=True)
        assert e1 < RTOL and e2 < RTOL
        del a


def test_bench_workload_apply(orc, h1pa, torch_cuda):
    """The benchmark's apply (unit cube 79^3, p = 4, vertex perturbation 0.2h
    with seed 1, sin/cos coefficient) against the oracle at full size."""
    M, torch = h1pa, torch_cuda
    n, p = 79, 4
    print(f"[scale] bench workload p={p} n={n}^3", flush=True)
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, 1)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = _form(M, fes, "diffusion", M.SinCosCoefficient())
    P = orc.Problem(n, p, perturb=0.2, seed=1, coeff="sincos")
    assert fes.GetNDofs() == P.ndofs
    x = np.random.default_rng(1).uniform(-1, 1, P.ndofs)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    a.Mult(xd, y)
    torch.cuda.synchronize()
    e = _rel(y.cpu().numpy(), P.mult(x))
    print(f"[scale] bench workload mult {e:.2e}", flush=True)
    assert e < RTOL
    assert np.array_equal(fes.GetElementRestriction().GatherMap().cpu().numpy(), P.gather_map)


def _cfg4_draw(count=16, seed=2024, max_dofs=400_000):
    """The first `count` problems of the batched workload's draw (the
    benchmark's generator, seed 2024) whose size the oracle's CPU PCG solves
    in seconds."""
    rng = np.random.default_rng(seed)
    ns, ps = rng.integers(4, 25, 512), rng.integers(2, 7, 512)
    out = []
    for i in range(count):
        n, p = int(ns[i]), int(ps[i])
        if (n * p + 1) ** 3 <= max_dofs:
            out.append((i, n, p))
    return out


def test_fourier_coefficient_pa_data(orc, h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    for (n, p, seed) in [(6, 3, 5), (4, 6, 77)]:
        mesh = M.Mesh.MakeCartesian3D(n, n, n)
        mesh.Perturb(0.15, seed)
        fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
        a = _form(M, fes, "diffusion", M.FourierCoefficient(seed))
        P = orc.Problem(n, p, perturb=0.15, seed=seed, coeff=("fourier", seed))
        dd, _ = a.PAData()
        torch.cuda.synchronize()
        assert _rel(dd.cpu().numpy(), P.Dd) < RTOL
        x = np.random.default_rng(seed).uniform(-1, 1, P.ndofs)
        y = torch.empty(P.ndofs, dtype=torch.float64, device="cuda")
        a.Mult(torch.from_numpy(x).cuda(), y)
        torch.cuda.synchronize()
        assert _rel(y.cpu().numpy(), P.mult(x)) < RTOL


def test_cfg4_batch_vs_oracle(orc, h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    cases = []
    for (i, n, p) in _cfg4_draw():
        print(f"[scale] cfg4 problem {i}: n={n} p={p}", flush=True)
        mesh = M.Mesh.MakeCartesian3D(n, n, n)
        mesh.Perturb(0.15, i)
        fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
        a = _form(M, fes, "diffusion", M.FourierCoefficient(i))
        P = orc.Problem(n, p, perturb=0.15, seed=i, coeff=("fourier", i))
        b = np.random.default_rng(1000 + i).uniform(-1, 1, P.ndofs)
        ess = fes.GetBoundaryTrueDofs()
        x0 = torch.zeros(P.ndofs, dtype=torch.float64, device="cuda")
        A, X, B = a.FormLinearSystem(ess, x0, torch.from_numpy(b).cuda())
        Xo, Bo = P.form_linear_system(np.zeros(P.ndofs), b)
        xo, st = P.pcg(Bo, rel_tol=1e-10, max_iter=20000)
        assert st["converged"]
        cases.append((i, A, X, B, P, Bo, xo, st, mesh, fes, a))
    assert len(cases) >= 10
    for prec in (M.PREC_FP64, M.PREC_BF16_PA):
        bs = M.BatchSolver(prec)
        Xs = {}
        for (i, A, X, B, *_r) in cases:
            Xs[i] = X.clone()
            bs.Add(A, B, Xs[i], rel_tol=1e-10, max_iter=20000, tag=i)
        assert bs.Run() == 0
        res = dict(bs.Poll())
        for (i, A, X, B, P, Bo, xo, st, *_r) in cases:
            x = Xs[i].cpu().numpy()
            r = Bo - P.constrained_mult(x)
            dinv = 1.0 / P.diagonal()
            dinv[P.ess] = 1.0
            meas = np.sqrt(r @ (dinv * r)) / np.sqrt(Bo @ (dinv * Bo))
            err = _rel(x, xo)
            print(f"[scale] cfg4 prec={prec} problem {i}: it {res[i]['iterations']} vs oracle "
                  f"{st['iterations']}, measure {meas:.2e}, diff {err:.2e}", flush=True)
            assert res[i]["converged"]
            if prec == M.PREC_FP64:
                assert abs(res[i]["iterations"] - st["iterations"]) <= 1
                assert err <= 1e-8
                assert meas <= 1.05e-10
            else:
                assert meas <= 1.05e-10
                assert err <= 1e-5
