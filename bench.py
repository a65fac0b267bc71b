"""Benchmark on one MI355X: PA diffusion apply (MDOF/s, the headline value),
plus the solve legs of the same metric (solves/s and p99 latency of
mixed-precision batched serving against plain FP64 GPU PCG at 1e-10).

Legs (all on by default; --legs selects):
  apply  configuration 2: unit cube, Cartesian hex mesh with interior vertices
         perturbed by U(-0.2h, 0.2h) (seed 1), H1 order p (default 4) with
         Q = p + 2 Gauss-Legendre points, k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y)
         sin(2 pi z) at the physical quadrature points, x ~ U[-1, 1].  A step is
         one y = A x (Operator::Mult: fused gather + element kernel, then the
         deterministic E->L scatter) with x, y and all operator data in HBM.
  solve  configuration 3: 48^3 mesh curved by x' = x + 0.05 sin(2 pi y)
         sin(2 pi z) (and cyclic), geometry of order p, per-element log-uniform
         k in [1e-2, 1e2], RHS U[-1, 1], homogeneous Dirichlet; FP64 Jacobi-PCG
         against the mixed solve (FP32 / BF16-stored PA inner CG, FP64 reliable
         updates) and CG preconditioned by p-multigrid (FP64 or BF16-PA cycle)
         at rel_tol 1e-10 (CGSolver's measure), p = 4 and 6; every solver is
         timed after a short warm-up solve of its own (first-use costs).
  batch  configuration 4: 512 problems, n ~ U[4, 24], p ~ U{2..6}, vertex
         perturbation 0.15h, k = exp(8 seeded Fourier modes), RHS U[-1, 1]:
         sequential FP64 CGSolver, FP64 batch and BF16 mixed batch; the same
         for 512 small problems (n ~ U[4, 8]).
  serve  configuration 5 (shortened arrival window): Poisson arrivals, n in
         [6, 32] with P(n) ~ 1/n, p ~ U{1..8}, 200 GB HBM budget; the native
         request loop mfem_amd/h1pa_serve: serial FP64 Jacobi-PCG (the
         baseline), continuous batching of BF16-PA mixed Jacobi-PCG solves in
         graph replays, and requests served side by side (4 solver streams)
         with CG preconditioned by p-multigrid, its cycle in FP64 or reading
         BF16 PA data.  Every mode stops at the same 1e-10 measure (the
         multigrid modes check each result on the baseline's Jacobi measure).
         The headline ratio is the BF16-cycle multigrid mode's.

  sweep  configuration 2 at about 4M and 32M DOFs for p = 1..8: PA diffusion
         and PA mass Mult, each against its own algorithmic bytes.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--p 4] [--dofs 32e6]
                    [--legs apply,sweep,solve,batch,serve] [--no-cpu]

Prints one JSON line.  `roofline` is for one whole Operator::Mult (both of
its kernels) against the fused apply's algorithmic bytes, timed with HIP
events on the launch stream, with the first kernel (the column sweep, or the
element kernel) as `roofline.element_kernel`; `cpu_baseline` times the CPU
oracle (C + OpenMP restatement of MFEM's PA path) on a bounded sample on the
host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "MDOF/s PA diffusion apply; solves/s + p99 latency vs FP64 PCG at 1e-10 rel res"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def n_for_dofs(p, target):
    """Elements per direction so that (n p + 1)^3 is closest to target."""
    best = None
    for n in range(1, 2000):
        d = (n * p + 1) ** 3
        if best is None or abs(d - target) < abs(best[1] - target):
            best = (n, d)
        if d > 2 * target:
            break
    return best[0]


def algorithmic_bytes(p, ne, ndofs):
    """Minimum HBM traffic of one fused L->L diffusion apply (FP64):
    PA data 6 Q^3 doubles + one int32 index per E-DOF per element, x read and
    y written once.  Element kernel alone: PA data + index + x read + E-vector
    write; scatter: offsets + indices + E-vector read + y write."""
    D, Q = p + 1, p + 2
    D3, Q3 = D ** 3, Q ** 3
    apply_ = ne * (6 * Q3 * 8 + D3 * 4) + 2 * ndofs * 8
    element = ne * (6 * Q3 * 8 + D3 * 4 + D3 * 8) + ndofs * 8
    scatter = (ndofs + 1) * 4 + ne * D3 * (4 + 8) + ndofs * 8
    return apply_, element, scatter


def mass_bytes(p, ne, ndofs):
    """Minimum HBM traffic of one fused L->L mass apply (FP64): PA data Q^3
    doubles + one int32 index per E-DOF per element, x read and y written."""
    D, Q = p + 1, p + 2
    return ne * (Q ** 3 * 8 + D ** 3 * 4) + 2 * ndofs * 8


# ---------------------------------------------------------------------------
# apply
# ---------------------------------------------------------------------------
def build_problem(M, torch, p, n, seed=1, integ="diffusion"):
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, seed)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    cls = M.DiffusionIntegrator if integ == "diffusion" else M.MassIntegrator
    a.AddDomainIntegrator(cls(M.SinCosCoefficient()))
    a.Assemble()
    N = fes.GetNDofs()
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    return mesh, fes, a, x, y


def run_apply(M, torch, p, n, steps, warmup, integ="diffusion"):
    """`steps` timed Operator::Mult calls; wall clock around the region
    (value) and HIP events on the launch stream around each kernel."""
    if integ == "diffusion":
        mesh, fes, a, x, y = build_problem(M, torch, p, n)
    else:
        mesh, fes, a, x, y = build_problem(M, torch, p, n, integ="mass")
    ne, N = fes.GetNE(), fes.GetNDofs()
    path = a.MultPath()
    for _ in range(warmup):
        a.Mult(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot_ms, el_ms, sc_ms = a.MultTimed(x, y, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / steps
    ev_ms = tot_ms / steps  # events: first launch start to last launch end
    if integ == "diffusion":
        b_apply, b_elem, b_scat = algorithmic_bytes(p, ne, N)
    else:
        b_apply = b_elem = mass_bytes(p, ne, N)
        b_scat = (N + 1) * 4 + ne * (p + 1) ** 3 * (4 + 8) + N * 8
    if path == "columns":
        b_elem = b_apply  # the column sweep moves the whole fused apply's bytes
    variant = a.Info()["variant"]
    del mesh, fes, a, x, y
    return dict(p=p, n=n, ne=ne, ndofs=N, path=path, ms_per_step=ms, event_ms=ev_ms,
                ms_element=el_ms, ms_scatter=sc_ms, mdofs=N / (ms * 1e-3) / 1e6,
                apply_gbs=b_apply / (ev_ms * 1e-3) / 1e9,
                element_gbs=b_elem / (el_ms * 1e-3) / 1e9,
                scatter_gbs=b_scat / (sc_ms * 1e-3) / 1e9 if sc_ms > 0 else None,
                bytes_apply=b_apply, bytes_element=b_elem, variant=variant)


def _event_ms(torch, fn, reps):
    """Average ms of fn() over reps launches, torch events on the current
    stream (the library launches on it)."""
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


# vector streams of one PCG iteration beyond the fused apply: the scatter's
# (d, q) dot reads d; the update reads x, d, r, q, dinv and writes x, r, z;
# the direction update reads z, d and writes d
PCG_VECTOR_STREAMS = 12


def run_path_kernels(M, torch, p, n, reps=20, pcg_iters=50):
    """The other kernels of the path on the same workload, each against its
    own algorithmic bytes: ElementRestriction gather (NE D^3 (8+4) + N 8),
    scatter (NE D^3 (8+4) + (N+1) 4 + N 8), PA diagonal (NE (6 Q^3 8 + D^3 4)
    + N 8) and one Jacobi-PCG iteration (the fused apply's bytes + 12 N 8)."""
    mesh, fes, a, x, y = build_problem(M, torch, p, n)
    ne, N = fes.GetNE(), fes.GetNDofs()
    D, Q = p + 1, p + 2
    D3, Q3 = D ** 3, Q ** 3
    R = fes.GetElementRestriction()
    xe = torch.empty(ne * D3, dtype=torch.float64, device="cuda")
    out = {}

    def rec(name, ms, nbytes, **kw):
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = dict(ms=round(ms, 4), bytes=int(nbytes), gbs=round(gbs, 1),
                         frac=round(gbs / HBM_PEAK_GBS, 4), **kw)

    rec("gather", _event_ms(torch, lambda: R.Mult(x, xe), reps), ne * D3 * 12 + N * 8)
    rec("scatter", _event_ms(torch, lambda: R.MultTranspose(xe, y), reps),
        ne * D3 * 12 + (N + 1) * 4 + N * 8)
    diag = torch.empty_like(x)
    rec("diagonal", _event_ms(torch, lambda: a.AssembleDiagonal(diag), reps),
        ne * (6 * Q3 * 8 + D3 * 4) + N * 8)
    ess = fes.GetBoundaryTrueDofs()
    A, X, B = a.FormLinearSystem(ess, torch.zeros_like(x), x)
    cg = M.CGSolver()
    cg.SetRelTol(0.0)
    cg.SetAbsTol(0.0)
    cg.SetMaxIter(pcg_iters)
    cg.SetOperator(A)
    cg.SetPreconditioner(M.OperatorJacobiSmoother(a, ess))
    X0 = X.clone()
    X.copy_(X0)
    cg.Mult(B, X)  # warm-up
    iters = []

    def solve():
        X.copy_(X0)
        cg.Mult(B, X)
        iters.append(cg.GetNumIterations())

    ms = _event_ms(torch, solve, 3)
    it = max(iters[-1], 1)
    b_apply = ne * (6 * Q3 * 8 + D3 * 4) + 2 * N * 8
    rec("pcg_iteration", ms / it, b_apply + PCG_VECTOR_STREAMS * N * 8, iterations=it,
        vector_streams=PCG_VECTOR_STREAMS)
    del mesh, fes, a, x, y, xe, diag, A, X, B, X0, cg
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, reps=10, warmup=3):
    """The CPU oracle (C + OpenMP restatement of MFEM's PA path, built -O3) on
    a bounded sample of the same workload: same p, perturbation and
    coefficient on a 2M-DOF mesh; MDOF/s of the L->L apply as the median of
    `reps` applies after `warmup`, on every host thread this job has
    (OMP_NUM_THREADS; 16 on the GPU box) and on one."""
    import numpy as np
    import oracle
    n = n_for_dofs(p, 2.0e6)
    P = oracle.Problem(n, p, perturb=0.2, seed=1, coeff="sincos")
    x = np.random.default_rng(1).uniform(-1, 1, P.ndofs)
    all_threads = oracle.num_threads()
    out = {}
    for threads in (all_threads, 1):
        oracle.set_num_threads(threads)
        for _ in range(warmup):
            P.mult(x)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            P.mult(x)
            ts.append(time.perf_counter() - t0)
        ts.sort()
        med = ts[len(ts) // 2] if len(ts) % 2 else 0.5 * (ts[len(ts) // 2 - 1] + ts[len(ts) // 2])
        out[threads] = P.ndofs / med / 1e6
    oracle.set_num_threads(all_threads)
    return dict(value=round(out[all_threads], 3), unit="MDOF/s", cores=all_threads,
                kind="port", value_1core=round(out[1], 3), cpu=cpu_model(),
                sample=f"oracle PA diffusion apply (FP64, C + OpenMP over elements), p={p}, "
                       f"n={n}^3 ({P.ndofs} DOFs): median of {reps} applies after {warmup} "
                       f"warm-ups, on {all_threads} threads and on 1")


def traffic_from_profiles(p, n, kernel, variant=None):
    """HBM bytes of one launch (kernel "mult_element+scatter": of one Mult,
    both kernels) measured with rocprofv3 PMC counters (FETCH_SIZE,
    WRITE_SIZE, scaled by the copy-kernel calibration) when
    profiles/traffic.json holds this workload, kernel and element-kernel
    variant; else None."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rows = json.load(f)
    except (OSError, ValueError):
        return None
    for r in rows:
        if r.get("p") == p and r.get("n") == n and r.get("kernel") == kernel and \
                (variant is None or r.get("variant_index") == variant):
            return r.get("bytes_per_launch")
    return None


# ---------------------------------------------------------------------------
# solve (configuration 3)
# ---------------------------------------------------------------------------
def mg_orders(p):
    """p-multigrid levels: the order halved (rounding up) down to 1."""
    orders = [p]
    while orders[-1] > 1:
        orders.append((orders[-1] + 1) // 2)
    return tuple(reversed(orders))


def run_solve(M, torch, p, n=48, seed=7):
    import numpy as np
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.SetCurvature(p, 0.05)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    ne, N = fes.GetNE(), fes.GetNDofs()
    g = torch.Generator(device="cuda").manual_seed(seed)
    k = torch.exp(torch.rand(ne, dtype=torch.float64, device="cuda", generator=g)
                  * (np.log(1e2) - np.log(1e-2)) + np.log(1e-2))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.PWConstCoefficient(k)))
    a.Assemble()
    b = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    ess = fes.GetBoundaryTrueDofs()
    x0 = torch.zeros(N, dtype=torch.float64, device="cuda")
    A, X, B = a.FormLinearSystem(ess, x0, b)
    y = torch.empty_like(B)

    def true_res(Xs):
        A.Mult(Xs, y)
        torch.cuda.synchronize()
        return (torch.linalg.norm(B - y) / torch.linalg.norm(B)).item()

    def fp64(max_iter):
        cg = M.CGSolver()
        cg.SetRelTol(1e-10)
        cg.SetMaxIter(max_iter)
        cg.SetOperator(A)
        cg.SetPreconditioner(M.OperatorJacobiSmoother(a, ess))
        Xs = X.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cg.Mult(B, Xs)
        torch.cuda.synchronize()
        return cg, Xs, time.perf_counter() - t0

    def mixed(prec, max_inner):
        ir = M.IRSolver(prec)
        ir.SetRelTol(1e-10)
        ir.SetDelta(0.1)
        ir.SetMaxIter(1000, max_inner)
        ir.SetOperator(A)
        Xs = X.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ir.Mult(B, Xs)
        torch.cuda.synchronize()
        return ir, Xs, time.perf_counter() - t0

    fp64(20)  # first-use costs
    cg, X64, t64 = fp64(100000)
    out = dict(p=p, n=n, dofs=N, fp64=dict(seconds=round(t64, 4), iterations=cg.GetNumIterations(),
                                          converged=cg.GetConverged(),
                                          true_rel_residual=true_res(X64)))
    # FP64 CG preconditioned by OperatorChebyshevSmoother (order 2, power-method
    # bound); set-up (diagonal + 10 power steps) timed separately
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cheb = M.OperatorChebyshevSmoother(A, None, ess, 2)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    cg = M.CGSolver()
    cg.SetRelTol(1e-10)
    cg.SetMaxIter(100000)
    cg.SetOperator(A)
    cg.SetPreconditioner(cheb)
    Xc = X.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cg.Mult(B, Xc)
    torch.cuda.synchronize()
    tc = time.perf_counter() - t0
    out["fp64_chebyshev2"] = dict(seconds=round(tc, 4), setup_seconds=round(t_setup, 4),
                                  iterations=cg.GetNumIterations(), converged=cg.GetConverged(),
                                  max_eig=cheb.GetMaxEigEstimate(),
                                  true_rel_residual=true_res(Xc))
    # FP64 CG preconditioned by p-multigrid (ex26's levels: orders halved down
    # to 1, Chebyshev(2) smoothing, coarse CG to 1e-2); set-up timed apart
    # (and the same with the cycle's applies reading BF16 PA data: mixed)
    orders = mg_orders(p)
    for name, cprec, cj in (("fp64_pmultigrid", M.PREC_FP64, False),
                            ("fp64_pmultigrid_bf16cycle", M.PREC_BF16_PA, False),
                            ("fp64_pmultigrid_bf16cycle_jacobi_coarse", M.PREC_BF16_PA, True)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        mg = M.PMultigrid(a, orders, cycle_precision=cprec, coarse_jacobi=cj)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        cg = M.CGSolver()
        cg.SetRelTol(1e-10)
        cg.SetOperator(A)
        cg.SetPreconditioner(mg)
        # first-use costs, as the FP64 baseline's warm-up: two iterations
        cg.SetMaxIter(2)
        cg.Mult(B, X.clone())
        coarse0 = mg.Info()["coarse_iterations"]
        cg.SetMaxIter(10000)
        Xm = X.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cg.Mult(B, Xm)
        torch.cuda.synchronize()
        tm = time.perf_counter() - t0
        info = mg.Info()
        info["coarse_iterations"] -= coarse0
        out[name] = dict(seconds=round(tm, 4), setup_seconds=round(t_setup, 4),
                         orders=list(orders), iterations=cg.GetNumIterations(),
                         converged=cg.GetConverged(),
                         coarse_iterations=info["coarse_iterations"],
                         true_rel_residual=true_res(Xm),
                         rel_diff_vs_fp64=(torch.linalg.norm(Xm - X64)
                                           / torch.linalg.norm(X64)).item(),
                         speedup_vs_fp64=round(t64 / tm, 3),
                         speedup_vs_fp64_with_setup=round(t64 / (tm + t_setup), 3))
        del mg, cg  # the hierarchy holds the form until it is destroyed
    for prec, name in ((M.PREC_FP32, "fp32"), (M.PREC_BF16_PA, "bf16")):
        mixed(prec, 20)
        ir, Xi, ti = mixed(prec, 100000)
        err = (torch.linalg.norm(Xi - X64) / torch.linalg.norm(X64)).item()
        out[name] = dict(seconds=round(ti, 4), iterations=ir.stats.inner_iterations,
                         refinements=ir.stats.outer_iterations, converged=ir.GetConverged(),
                         true_rel_residual=true_res(Xi), rel_diff_vs_fp64=err,
                         speedup_vs_fp64=round(t64 / ti, 3))
    return out


# ---------------------------------------------------------------------------
# batch (configuration 4)
# ---------------------------------------------------------------------------
def run_batch(M, torch, nprob=512, seed=2024, nmin=4, nmax=24):
    """Configuration 4: nprob solves, n uniform in [nmin, nmax], p in 2..6,
    Fourier coefficient, rel_tol 1e-10: sequential FP64 CGSolver, the FP64
    batch and the BF16-PA mixed batch (graph replays)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns, ps = rng.integers(nmin, nmax + 1, nprob), rng.integers(2, 7, nprob)
    probs = []
    t0 = time.perf_counter()
    for i in range(nprob):
        n, p = int(ns[i]), int(ps[i])
        mesh = M.Mesh.MakeCartesian3D(n, n, n)
        mesh.Perturb(0.15, i)
        fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
        a = M.BilinearForm(fes)
        a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
        a.AddDomainIntegrator(M.DiffusionIntegrator(M.FourierCoefficient(i)))
        a.Assemble()
        N = fes.GetNDofs()
        b = torch.empty(N, dtype=torch.float64, device="cuda")
        M.Randomize(b, 1000 + i, -1.0, 1.0)
        ess = fes.GetBoundaryTrueDofs()
        A, X, B = a.FormLinearSystem(ess, torch.zeros(N, dtype=torch.float64, device="cuda"), b)
        probs.append((mesh, fes, a, A, X, B, N))
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    total = sum(pr[6] for pr in probs)
    log(f"batch: {nprob} problems, {total / 1e6:.1f}M DOFs set up in {setup:.1f} s")

    def sequential():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its, xs = [], []
        for pr in probs:
            cg = M.CGSolver()
            cg.SetRelTol(1e-10)
            cg.SetMaxIter(100000)
            cg.SetOperator(pr[3])
            cg.SetPreconditioner(M.OperatorJacobiSmoother(pr[2], pr[3].ess))
            x = pr[4].clone()
            cg.Mult(pr[5], x)
            its.append(cg.GetNumIterations())
            xs.append(x)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, its, xs

    def batched(prec):
        # timed from the first Add: each Add does the per-solve set-up
        # (Jacobi diagonal, constraint mask, shadow PA data) that the
        # sequential CGSolver::Mult does inside its own call
        bs = M.BatchSolver(prec, iters_per_replay=8)
        xs = [pr[4].clone() for pr in probs]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, pr in enumerate(probs):
            bs.Add(pr[3], pr[5], xs[i], rel_tol=1e-10, max_iter=100000, tag=i)
        left = bs.Run()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        res = dict(bs.Poll())
        assert left == 0
        return t, [res[i] for i in range(nprob)], xs, bs.Info()

    t_seq, its, x_seq = sequential()
    t_b64, r64, x_b64, _ = batched(M.PREC_FP64)
    t_bm, rm, x_bm, info = batched(M.PREC_BF16_PA)
    diff = max((torch.linalg.norm(a_ - b_) / torch.linalg.norm(b_)).item()
               for a_, b_ in zip(x_bm, x_seq))
    return dict(
        problems=nprob, n_range=[nmin, nmax], total_dofs=total, setup_seconds=round(setup, 2),
        sequential_fp64=dict(seconds=round(t_seq, 3), solves_per_s=round(nprob / t_seq, 2),
                             iterations_mean=float(np.mean(its)), iterations_max=int(max(its))),
        batch_fp64=dict(seconds=round(t_b64, 3), solves_per_s=round(nprob / t_b64, 2),
                        converged=sum(r["converged"] for r in r64)),
        batch_bf16=dict(seconds=round(t_bm, 3), solves_per_s=round(nprob / t_bm, 2),
                        converged=sum(r["converged"] for r in rm),
                        iterations_mean=float(np.mean([r["iterations"] for r in rm])),
                        max_rel_diff_vs_fp64=diff, replays=info["replays"],
                        graph_captures=info["captures"]),
        mixed_batch_vs_sequential_fp64=round(t_seq / t_bm, 3))


# ---------------------------------------------------------------------------
# serve (configuration 5, shortened window)
# ---------------------------------------------------------------------------
def run_serve(rate, seconds, seed=1):
    exe = os.path.join(HERE, "mfem_amd", "h1pa_serve")
    if not os.path.exists(exe):
        raise RuntimeError("mfem_amd/h1pa_serve is not built (run __graft_entry__.build())")
    out = {"rate_per_s": rate, "window_s": seconds}
    for mode, prec in (("serial", "fp64"), ("batch", "bf16"), ("mg", "bf16"), ("mg", "fp64")):
        cmd = [exe, "--mode", mode, "--precision", prec, "--rate", str(rate), "--seconds",
               str(seconds), "--seed", str(seed)]
        log("serve: " + " ".join(cmd[1:]))
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError(f"h1pa_serve failed: {r.stderr[-2000:]}")
        out[f"{mode}_{prec}"] = json.loads(r.stdout.strip().splitlines()[-1])
    s = out["serial_fp64"]["solves_per_s"]
    for k in ("batch_bf16", "mg_bf16", "mg_fp64"):
        out[f"{k}_ratio"] = round(out[k]["solves_per_s"] / s, 3)
    # the headline serving mode: mixed precision (BF16-PA cycle in an FP64 CG)
    # with p-multigrid, requests served side by side
    out["solves_per_s_ratio"] = out["mg_bf16_ratio"]
    return out


# Arrival rates of configuration 5 with windows of about 200 requests each:
# the mix is ~10 s of GPU work per second of arrivals at 200/s, so at every
# rate the box is overloaded and a longer window only lengthens the queue.
SERVE_RATES = ((200.0, 1.0), (1000.0, 0.2), (5000.0, 0.04))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--dofs", type=float, default=32e6)
    ap.add_argument("--legs", default="apply,sweep,solve,batch,serve")
    ap.add_argument("--sweep-dofs", default="4e6,32e6",
                    help="DOF targets of the p = 1..8 sweep (PA diffusion and PA mass)")
    ap.add_argument("--serve-rates", default=",".join(f"{r:g}:{w:g}" for r, w in SERVE_RATES),
                    help="rate:window pairs of the serving leg")
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device copy first (scales rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(one process, one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    legs = set(args.legs.split(","))

    import torch
    import mfem_amd as M
    M.lib()
    torch.cuda.init()
    if args.calib_copy:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    n = n_for_dofs(args.p, args.dofs)
    log(f"apply: PA diffusion p={args.p} n={n}^3")
    r = run_apply(M, torch, args.p, n, args.steps, args.warmup)
    log(str(r))
    torch.cuda.empty_cache()
    path_kernels = {}
    if "apply" in legs:
        log("path kernels: gather, scatter, diagonal, PCG iteration")
        path_kernels = run_path_kernels(M, torch, args.p, n)
        log(str(path_kernels))
        torch.cuda.empty_cache()
    sweep = []
    if "sweep" in legs:
        for target in (float(v) for v in args.sweep_dofs.split(",") if v):
            for p in range(1, 9):
                ns = n_for_dofs(p, target)
                row = dict(p=p, n=ns, target_dofs=target)
                for integ in ("diffusion", "mass"):
                    log(f"sweep {integ} p={p} n={ns}^3")
                    s_ = run_apply(M, torch, p, ns, 10, 3, integ=integ)
                    row["dofs"] = s_["ndofs"]
                    row[integ] = dict(path=s_["path"], ms=round(s_["event_ms"], 4),
                                      mdofs=round(s_["ndofs"] / (s_["event_ms"] * 1e-3) / 1e6, 1),
                                      gbs=round(s_["apply_gbs"], 1),
                                      frac=round(s_["apply_gbs"] / HBM_PEAK_GBS, 4))
                    torch.cuda.empty_cache()
                sweep.append(row)

    extra = {}
    if "solve" in legs:
        extra["solve"] = []
        for p in (4, 6):
            log(f"solve: configuration 3, p={p}")
            extra["solve"].append(run_solve(M, torch, p))
            torch.cuda.empty_cache()
    if "batch" in legs:
        log(f"batch: configuration 4, {args.batch_problems} problems")
        extra["batch"] = run_batch(M, torch, args.batch_problems)
        torch.cuda.empty_cache()
        log(f"batch: small problems (n in [4, 8]), {args.batch_problems} problems")
        extra["batch_small"] = run_batch(M, torch, args.batch_problems, seed=2025, nmin=4, nmax=8)
        torch.cuda.empty_cache()
    if "serve" in legs:
        extra["serve"] = []
        for pair in args.serve_rates.split(","):
            rate, window = (float(v) for v in pair.split(":"))
            extra["serve"].append(run_serve(rate, window))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle)")
        cpu = cpu_baseline(args.p)

    mult_gbs = r["apply_gbs"]
    line = {
        "metric": METRIC,
        "value": round(r["mdofs"], 2),
        "unit": "MDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"PA diffusion apply (Operator::Mult, L-vector to L-vector), unit cube "
                        f"{n}^3 hex, vertex perturbation 0.2h, H1 p={args.p} Q={args.p + 2}, "
                        f"k=1+0.5sin(2pi x)cos(2pi y)sin(2pi z)",
            "p": args.p, "q1d": args.p + 2, "elements": r["ne"], "dofs": r["ndofs"],
            "mult_path": r["path"], "element_variant": r["variant"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(mult_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(mult_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(args.p, n, "mult_" + r["path"], r["variant"]),
            "kernel": ("k_diffusion_columns + k_columns_interface (one Operator::Mult)"
                       if r["path"] == "columns" else
                       "k_diffusion_apply + k_restrict_mult_t (one Operator::Mult)"),
            "ms": round(r["event_ms"], 4),
            "bytes_per_launch": r["bytes_apply"],
            "bytes_formula": "NE (6 Q^3 8 + D^3 4) + 2 N 8",
            "element_kernel": {
                "kernel": ("k_diffusion_columns" if r["path"] == "columns" else
                           f"k_diffusion_apply<double,double,{args.p + 1},{args.p + 2},...>"),
                "ms": round(r["ms_element"], 4),
                "bytes": r["bytes_element"],
                "gbs": round(r["element_gbs"], 1),
                "frac": round(r["element_gbs"] / HBM_PEAK_GBS, 4),
                "second_kernel_ms": round(r["ms_scatter"], 4),
            },
        },
        "kernels": path_kernels,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    if "serve" in extra and extra["serve"]:
        s0 = extra["serve"][0]
        line["serving_mode"] = ("mg bf16: CG preconditioned by p-multigrid whose cycle reads "
                                "BF16 PA data (FP64 vectors, FP64 outer CG), requests solved "
                                "side by side on 4 streams")
        line["solves_per_s"] = s0["mg_bf16"]["solves_per_s"]
        line["p99_latency_s"] = s0["mg_bf16"]["latency_p99_s"]
        line["max_jacobi_rel_residual"] = s0["mg_bf16"]["max_jacobi_rel_residual"]
        line["fp64_pcg_solves_per_s"] = s0["serial_fp64"]["solves_per_s"]
        line["fp64_pcg_p99_latency_s"] = s0["serial_fp64"]["latency_p99_s"]
        line["solves_per_s_ratio"] = s0["solves_per_s_ratio"]
        line["batch_bf16_jacobi_solves_per_s_ratio"] = s0["batch_bf16_ratio"]
    if sweep:
        line["sweep"] = sweep
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
