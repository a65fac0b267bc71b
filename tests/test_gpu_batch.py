"""GPU: batched solves (no MFEM counterpart) against single solves and the
oracle.

Bars, written per test:
  - FP64 batch vs the single-solve CGSolver path: iteration counts within +-1,
    solutions within 1e-9 relative (both at rel_tol 1e-12).
  - Alone vs in a mix, graph vs plain launches, added mid-run: bitwise equal
    solutions and equal iteration counts (each solve's arithmetic and
    reduction order do not depend on the other solves).
  - Mixed batch: CGSolver's measure on the FP64 true residual <= 1e-10 and the
    solution within 1e-6 relative of the oracle's FP64 PCG solution.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


class Prob:
    """cfg4-style problem: n^3 mesh, vertex perturbation 0.15h, k = exp(8
    Fourier modes), RHS uniform[-1, 1], homogeneous Dirichlet."""

    def __init__(self, M, torch, n, p, seed):
        print(f"[prob] n={n} p={p} seed={seed}", flush=True)
        self.n_, self.p, self.seed = n, p, seed
        self.mesh = M.Mesh.MakeCartesian3D(n, n, n)
        self.mesh.Perturb(0.15, seed)
        self.fes = M.FiniteElementSpace(self.mesh, M.H1_FECollection(p, 3))
        self.a = M.BilinearForm(self.fes)
        self.a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
        self.a.AddDomainIntegrator(M.DiffusionIntegrator(M.FourierCoefficient(seed)))
        self.a.Assemble()
        N = self.fes.GetNDofs()
        self.N = N
        g = torch.Generator(device="cuda").manual_seed(1000 + seed)
        b = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
        self.ess = self.fes.GetBoundaryTrueDofs()
        x0 = torch.zeros(N, dtype=torch.float64, device="cuda")
        self.A, self.X0, self.B = self.a.FormLinearSystem(self.ess, x0, b)
        self.torch = torch

    def fresh_x(self):
        return self.X0.clone()


SPECS = [(4, 2, 1), (6, 3, 2), (3, 5, 3), (5, 4, 4), (3, 6, 5), (7, 2, 6), (4, 3, 7)]


def _run_batch(M, probs, precision, use_graph=True, k=8, rel_tol=1e-12, add_later=()):
    """Solve probs together (and add_later after two replays); returns
    {seed: (x, result)}."""
    bs = M.BatchSolver(precision, iters_per_replay=k, use_graph=use_graph)
    xs = {}
    for P in probs:
        xs[P.seed] = P.fresh_x()
        bs.Add(P.A, P.B, xs[P.seed], rel_tol=rel_tol, max_iter=20000, tag=P.seed)
    res = {}
    if add_later:
        bs.Run(2)
        for tag, r in bs.Poll():
            res[tag] = r
        for P in add_later:
            xs[P.seed] = P.fresh_x()
            bs.Add(P.A, P.B, xs[P.seed], rel_tol=rel_tol, max_iter=20000, tag=P.seed)
    left = bs.Run()
    assert left == 0
    for tag, r in bs.Poll():
        res[tag] = r
    info = bs.Info()
    assert info["live"] == 0
    return {s: (xs[s].cpu().numpy(), res[s]) for s in xs}


@pytest.fixture(scope="module")
def probs(h1pa, torch_cuda):
    return [Prob(h1pa, torch_cuda, *s) for s in SPECS]


def test_fp64_batch_matches_single_solves(h1pa, torch_cuda, probs):
    out = _run_batch(h1pa, probs, h1pa.PREC_FP64)
    for P in probs:
        cg = h1pa.CGSolver()
        cg.SetRelTol(1e-12)
        cg.SetMaxIter(20000)
        cg.SetOperator(P.A)
        cg.SetPreconditioner(h1pa.OperatorJacobiSmoother(P.a, P.ess))
        X = P.fresh_x()
        cg.Mult(P.B, X)
        x, r = out[P.seed]
        print(f"[batch64] n={P.n_} p={P.p} iters batch={r['iterations']} single="
              f"{cg.GetNumIterations()} diff={_rel(x, X.cpu().numpy()):.2e}", flush=True)
        assert r["converged"] and cg.GetConverged()
        assert abs(r["iterations"] - cg.GetNumIterations()) <= 1
        assert _rel(x, X.cpu().numpy()) < 1e-9


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_alone_equals_mix_bitwise(h1pa, torch_cuda, probs, prec):
    P = {"fp64": h1pa.PREC_FP64, "fp32": h1pa.PREC_FP32, "bf16": h1pa.PREC_BF16_PA}[prec]
    rel = 1e-12 if prec == "fp64" else 1e-10
    mix = _run_batch(h1pa, probs, P, rel_tol=rel)
    for Q in probs[:3]:
        alone = _run_batch(h1pa, [Q], P, rel_tol=rel)
        xa, ra = alone[Q.seed]
        xm, rm = mix[Q.seed]
        print(f"[mix] {prec} seed={Q.seed} iters alone={ra['iterations']} mix={rm['iterations']}",
              flush=True)
        assert ra == rm
        assert np.array_equal(xa, xm)


def test_graph_equals_plain_launches(h1pa, torch_cuda, probs):
    g = _run_batch(h1pa, probs[:4], h1pa.PREC_BF16_PA, use_graph=True)
    p = _run_batch(h1pa, probs[:4], h1pa.PREC_BF16_PA, use_graph=False)
    for s in g:
        assert g[s][1] == p[s][1]
        assert np.array_equal(g[s][0], p[s][0])


@pytest.mark.parametrize("prec", ["fp64", "bf16"])
def test_continuous_admission(h1pa, torch_cuda, probs, prec):
    """Solves added while others are on the GPU (and into retired slots) give
    the same bits as alone."""
    P = h1pa.PREC_FP64 if prec == "fp64" else h1pa.PREC_BF16_PA
    rel = 1e-12 if prec == "fp64" else 1e-10
    out = _run_batch(h1pa, probs[:4], P, rel_tol=rel, add_later=probs[4:])
    for Q in probs[4:]:
        alone = _run_batch(h1pa, [Q], P, rel_tol=rel)
        assert alone[Q.seed][1] == out[Q.seed][1]
        assert np.array_equal(alone[Q.seed][0], out[Q.seed][0])


def test_mixed_batch_vs_oracle(orc, h1pa, torch_cuda):
    """Mixed batch on perturbed meshes with a per-element log-uniform
    coefficient, against the oracle's FP64 PCG."""
    torch = torch_cuda
    cases = []
    for (n, p, seed) in [(4, 3, 11), (3, 5, 12), (5, 2, 13)]:
        ne = n ** 3
        k = np.exp(np.random.default_rng(seed).uniform(np.log(1e-2), np.log(1e2), ne))
        m = h1pa.Mesh.MakeCartesian3D(n, n, n)
        m.Perturb(0.15, seed)
        fes = h1pa.FiniteElementSpace(m, h1pa.H1_FECollection(p, 3))
        a = h1pa.BilinearForm(fes)
        a.SetAssemblyLevel(h1pa.AssemblyLevel.PARTIAL)
        kd = torch.from_numpy(k).cuda()
        a.AddDomainIntegrator(h1pa.DiffusionIntegrator(h1pa.PWConstCoefficient(kd)))
        a.Assemble()
        Po = orc.Problem(n, p, perturb=0.15, seed=seed, coeff=k)
        b = np.random.default_rng(seed + 1).uniform(-1, 1, Po.ndofs)
        ess = fes.GetBoundaryTrueDofs()
        x0 = torch.zeros(Po.ndofs, dtype=torch.float64, device="cuda")
        A, X, B = a.FormLinearSystem(ess, x0, torch.from_numpy(b).cuda())
        Xo, Bo = Po.form_linear_system(np.zeros(Po.ndofs), b)
        xo, st = Po.pcg(Bo, rel_tol=1e-13, max_iter=20000)
        cases.append((seed, A, X, B, Po, Bo, xo, kd, m, fes, a))
    for prec in (h1pa.PREC_FP32, h1pa.PREC_BF16_PA):
        bs = h1pa.BatchSolver(prec)
        Xs = {}
        for (seed, A, X, B, Po, Bo, xo, *_rest) in cases:
            Xs[seed] = X.clone()
            bs.Add(A, B, Xs[seed], rel_tol=1e-10, tag=seed)
        assert bs.Run() == 0
        res = dict(bs.Poll())
        for (seed, A, X, B, Po, Bo, xo, *_rest) in cases:
            x = Xs[seed].cpu().numpy()
            r = Bo - Po.constrained_mult(x)
            dinv = 1.0 / Po.diagonal()
            dinv[Po.ess] = 1.0
            meas = np.sqrt(r @ (dinv * r)) / np.sqrt(Bo @ (dinv * Bo))
            err = _rel(x, xo)
            print(f"[batchmix] prec={prec} seed={seed} {res[seed]} oracle_meas={meas:.2e} "
                  f"err={err:.2e}", flush=True)
            assert res[seed]["converged"]
            assert meas <= 1.05e-10
            assert err <= 1e-6


def test_add_while_replay_in_flight(h1pa, torch_cuda, probs):
    """Solves added while a replay runs (Launch ... Add ... Wait) join at the
    next launch and give the same bits as alone."""
    P = h1pa.PREC_BF16_PA
    first, later = probs[:3], probs[3:6]
    bs = h1pa.BatchSolver(P, iters_per_replay=4)
    xs = {}
    for Q in first:
        xs[Q.seed] = Q.fresh_x()
        bs.Add(Q.A, Q.B, xs[Q.seed], rel_tol=1e-10, tag=Q.seed)
    assert bs.Launch()
    for Q in later:  # replay in flight
        xs[Q.seed] = Q.fresh_x()
        bs.Add(Q.A, Q.B, xs[Q.seed], rel_tol=1e-10, tag=Q.seed)
    res = {}
    while True:
        bs.Wait()
        res.update(dict(bs.Poll()))
        if not bs.Launch():
            break
    assert sorted(res) == sorted(xs)
    for Q in later:
        alone = _run_batch(h1pa, [Q], P, rel_tol=1e-10, k=4)
        assert alone[Q.seed][1] == res[Q.seed]
        assert np.array_equal(alone[Q.seed][0], xs[Q.seed].cpu().numpy())


def test_batch_query_loop_matches_run(h1pa, torch_cuda):
    """Replays driven by Launch / Idle (the non-blocking h1pa_batch_query) /
    Wait, as the serving loop drives them, give bitwise Run's results; Idle
    is true whenever nothing is in flight."""
    import time
    M = h1pa
    P = Prob(M, torch_cuda, 4, 3, 21)
    out = []
    for mode in ("run", "query"):
        bs = M.BatchSolver(M.PREC_BF16_PA, iters_per_replay=8)
        X = P.fresh_x()
        bs.Add(P.A, P.B, X, rel_tol=1e-10, max_iter=20000, tag=7)
        assert bs.Idle()
        if mode == "run":
            assert bs.Run() == 0
        else:
            t0 = time.time()
            while bs.Launch():
                while not bs.Idle():
                    assert time.time() - t0 < 120, "replay did not finish"
                    time.sleep(1e-4)
                bs.Wait()
                assert bs.Idle()
        res = dict(bs.Poll())
        assert res[7]["converged"]
        out.append((X.cpu().numpy(), res[7]["iterations"]))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_form_in_live_batch_is_protected(h1pa, torch_cuda):
    """A live batch solve holds raw pointers to its form's PA data: until Poll
    retires it, re-assembling the form, building its reduced-precision shadow
    of the other kind (another batch's precision, MultLowPrecision) and
    destroying it are refused; after the poll they work again."""
    M = h1pa
    P = Prob(M, torch_cuda, 4, 3, 31)
    bs = M.BatchSolver(M.PREC_BF16_PA, iters_per_replay=8)
    X = P.fresh_x()
    bs.Add(P.A, P.B, X, rel_tol=1e-10, max_iter=20000, tag=1)
    with pytest.raises(M.H1paError) as ei:
        P.a.Assemble()
    assert "batch" in str(ei.value)
    other = M.BatchSolver(M.PREC_FP32, iters_per_replay=8)
    with pytest.raises(M.H1paError):
        other.Add(P.A, P.B, P.fresh_x(), rel_tol=1e-10, max_iter=20000, tag=2)
    x32 = torch_cuda.zeros(P.N, dtype=torch_cuda.float32, device="cuda")
    y32 = torch_cuda.empty_like(x32)
    with pytest.raises(M.H1paError):
        P.a.MultLowPrecision(x32, y32, M.PREC_FP32)
    f = P.a._single()
    assert M.lib().h1pa_form_destroy(f._h) == M.H1PA_EINVAL
    # the same precision shares the shadow: allowed while the solve is live
    P.a.MultLowPrecision(x32, y32, M.PREC_BF16_PA)
    assert bs.Run() == 0
    res = dict(bs.Poll())
    assert res[1]["converged"]
    P.a.Assemble()
    P.a.MultLowPrecision(x32, y32, M.PREC_FP32)
    torch_cuda.cuda.synchronize()
