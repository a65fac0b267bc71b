"""Benchmark: PA diffusion operator apply (MDOF/s) on one MI355X.

Workload (configuration 2 of the sweep): unit cube, Cartesian hex mesh with
interior vertices perturbed by U(-0.2h, 0.2h) (seed 1), H1 order p with
Q = p + 2 Gauss-Legendre points, diffusion coefficient
k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y) sin(2 pi z) at the physical quadrature
points, input x ~ U[-1, 1] over the true DOFs.  A step is one y = A x
(Operator::Mult: fused gather + element kernel, then the deterministic E->L
scatter) with x, y and all operator data resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--p 4] [--dofs 32e6]
                    [--sweep] [--no-cpu]

Prints one JSON line (rank 0).  `roofline` is for the dominant kernel (the
element kernel); `cpu_baseline` times the CPU oracle (C + OpenMP restatement
of MFEM's PA path) on a bounded sample on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


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


def build_problem(M, torch, p, n, seed=1):
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, seed)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient()))
    a.Assemble()
    N = fes.GetNDofs()
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    y = torch.empty_like(x)
    return mesh, fes, a, x, y


def run_gpu(M, torch, p, n, steps, warmup):
    mesh, fes, a, x, y = build_problem(M, torch, p, n)
    ne, N = fes.GetNE(), fes.GetNDofs()
    for _ in range(warmup):
        a.Mult(x, y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tot_ms, el_ms, sc_ms = a.MultTimed(x, y, steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall * 1e3 / steps
    b_apply, b_elem, b_scat = algorithmic_bytes(p, ne, N)
    return dict(p=p, n=n, ne=ne, ndofs=N, ms_per_step=ms, event_ms_total=tot_ms / steps,
                ms_element=el_ms, ms_scatter=sc_ms, mdofs=N / (ms * 1e-3) / 1e6,
                apply_gbs=b_apply / (ms * 1e-3) / 1e9, element_gbs=b_elem / (el_ms * 1e-3) / 1e9,
                scatter_gbs=b_scat / (sc_ms * 1e-3) / 1e9, bytes_apply=b_apply,
                bytes_element=b_elem)


def cpu_baseline(p, target_seconds=15.0):
    """The CPU oracle (C + OpenMP restatement of MFEM's PA path) on a bounded
    sample of the same workload: same p, perturbation and coefficient, on a
    smaller mesh; MDOF/s of the L->L apply."""
    import numpy as np
    import oracle
    n = n_for_dofs(p, 2.0e6)
    P = oracle.Problem(n, p, perturb=0.2, seed=1, coeff="sincos")
    x = np.random.default_rng(1).uniform(-1, 1, P.ndofs)
    P.mult(x)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        P.mult(x)
        reps += 1
        el = time.perf_counter() - t0
        if el > target_seconds or reps >= 50:
            break
    return dict(value=P.ndofs * reps / el / 1e6, unit="MDOF/s", cores=oracle.num_threads(),
                kind="port",
                sample=f"oracle PA diffusion apply, p={p}, n={n}^3 ({P.ndofs} DOFs), "
                       f"{reps} applies in {el:.1f} s, OpenMP over elements")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--dofs", type=float, default=32e6)
    ap.add_argument("--sweep", action="store_true", help="also run p = 1..8 at --sweep-dofs")
    ap.add_argument("--sweep-dofs", type=float, default=4e6)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 1 GiB device copy first (scales rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(one process, one MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

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
    print(f"[bench] PA diffusion apply p={args.p} n={n}^3", file=sys.stderr, flush=True)
    r = run_gpu(M, torch, args.p, n, args.steps, args.warmup)
    print(f"[bench] {r}", file=sys.stderr, flush=True)
    sweep = []
    if args.sweep:
        for p in range(1, 9):
            ns = n_for_dofs(p, args.sweep_dofs)
            print(f"[bench] sweep p={p} n={ns}^3", file=sys.stderr, flush=True)
            s = run_gpu(M, torch, p, ns, max(10, args.steps // 2), 3)
            sweep.append({k: s[k] for k in ("p", "n", "ndofs", "mdofs", "apply_gbs",
                                             "element_gbs", "ms_per_step")})
            torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.p)

    line = {
        "metric": "MDOF/s PA diffusion apply",
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
            "workload": f"PA diffusion apply, unit cube {n}^3 hex, vertex perturbation 0.2h, "
                        f"H1 p={args.p} Q={args.p + 2}, k=1+0.5sin(2pi x)cos(2pi y)sin(2pi z)",
            "p": args.p, "q1d": args.p + 2, "elements": r["ne"], "dofs": r["ndofs"],
            "apply_gbs_algorithmic": round(r["apply_gbs"], 1),
            "apply_roofline_frac": round(r["apply_gbs"] / HBM_PEAK_GBS, 4),
            "scatter_ms": round(r["ms_scatter"], 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(r["element_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r["element_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": f"k_diffusion_apply<{args.p + 1},{args.p + 2}>",
            "kernel_ms": round(r["ms_element"], 4),
            "bytes_per_launch": r["bytes_element"],
        },
        "cpu_baseline": cpu,
    }
    if sweep:
        line["sweep"] = sweep
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
