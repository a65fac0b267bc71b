"""GPU: every element-kernel variant of the diffusion apply computes the same
bits.

The variants (elements per block, line-loop unroll, PA-data pipeline depth,
register cap) differ only in scheduling; the autotuner picks one per (p, Q)
by timing at sizes no other test runs, so each is forced here in turn with
H1PA_DIFF_VARIANT (read when a form is assembled, and when its reduced-
precision PA data is made).

Bars: FP64 applies bitwise equal across variants 0..8 and within 1e-13
relative of the oracle; FP32 and BF16-PA applies bitwise equal across
variants 0, 1, 4, 5, 6, 7, 8; a FP64 and a mixed batch give bitwise the same
solutions and iteration counts under every variant.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

V64 = list(range(9))
V32 = [0, 1, 4, 5, 6, 7, 8]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _form(M, mesh, p, q1d, integrators=("diffusion",)):
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    for it in integrators:
        cls = M.DiffusionIntegrator if it == "diffusion" else M.MassIntegrator
        a.AddDomainIntegrator(cls(M.SinCosCoefficient(), q1d))
    a.Assemble()
    return fes, a


def _dims(p):
    return (3, 2, 3) if p <= 5 else (2, 2, 2)


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("dq", [1, 2])
def test_fp64_variants_bitwise(orc, h1pa, torch_cuda, monkeypatch, p, dq):
    torch, M = torch_cuda, h1pa
    # Mult of a diffusion form at q1d = p + 2 on a Cartesian mesh takes the
    # fused column apply, which has no element-kernel variants: keep Mult on
    # the element kernels here
    monkeypatch.setenv("H1PA_COLUMNS", "0")
    q1d = p + dq
    dims = _dims(p)
    P = orc.Problem(None, p, Q=q1d, perturb=0.2, seed=3, coeff="sincos", dims=dims)
    x = np.random.default_rng(p).uniform(-1, 1, P.ndofs)
    yo = P.mult(x)
    mesh = M.Mesh.MakeCartesian3D(*dims)
    mesh.Perturb(0.2, 3)
    xd = torch.from_numpy(x).cuda()
    ref = None
    for v in V64:
        print(f"[variant] fp64 p={p} q1d={q1d} v={v}", flush=True)
        monkeypatch.setenv("H1PA_DIFF_VARIANT", str(v))
        fes, a = _form(M, mesh, p, q1d)
        assert a.Info()["variant"] == v
        y = torch.empty(P.ndofs, dtype=torch.float64, device="cuda")
        a.Mult(xd, y)
        torch.cuda.synchronize()
        yh = y.cpu().numpy()
        assert _rel(yh, yo) < 1e-13
        if ref is None:
            ref = yh
        else:
            assert np.array_equal(yh, ref), f"variant {v} differs from variant 0"


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_low_precision_variants_bitwise(h1pa, torch_cuda, monkeypatch, p, prec):
    torch, M = torch_cuda, h1pa
    kind = M.PREC_FP32 if prec == "fp32" else M.PREC_BF16_PA
    mesh = M.Mesh.MakeCartesian3D(*_dims(p))
    mesh.Perturb(0.2, 4)
    ref = None
    for v in V32:
        print(f"[variant] {prec} p={p} v={v}", flush=True)
        monkeypatch.setenv("H1PA_DIFF_VARIANT", str(v))
        fes, a = _form(M, mesh, p, p + 2)
        n = fes.GetNDofs()
        x = torch.from_numpy(np.random.default_rng(p).uniform(-1, 1, n).astype(np.float32)).cuda()
        y = torch.empty(n, dtype=torch.float32, device="cuda")
        a.MultLowPrecision(x, y, kind)
        torch.cuda.synchronize()
        yh = y.cpu().numpy()
        assert np.all(np.isfinite(yh))
        if ref is None:
            ref = yh
        else:
            assert np.array_equal(yh, ref), f"variant {v} differs from variant 0"


def test_variant_out_of_range_is_an_error(h1pa, torch_cuda, monkeypatch):
    M = h1pa
    mesh = M.Mesh.MakeCartesian3D(2, 2, 2)
    monkeypatch.setenv("H1PA_DIFF_VARIANT", "9")  # one past the last
    with pytest.raises(M.H1paError):
        _form(M, mesh, 2, 4)
    monkeypatch.setenv("H1PA_DIFF_VARIANT", "2")  # FP64-only variant
    fes, a = _form(M, mesh, 2, 4)
    x = torch_cuda.zeros(fes.GetNDofs(), dtype=torch_cuda.float32, device="cuda")
    y = torch_cuda.empty_like(x)
    with pytest.raises(M.H1paError):
        a.MultLowPrecision(x, y, M.PREC_FP32)


@pytest.mark.parametrize("precision", ["fp64", "bf16"])
def test_batch_variants_bitwise(h1pa, torch_cuda, monkeypatch, precision):
    torch, M = torch_cuda, h1pa
    prec = M.PREC_FP64 if precision == "fp64" else M.PREC_BF16_PA
    specs = [(3, 2), (2, 4), (2, 6), (1, 8)]  # (n, p)
    ref = None
    for v in (V64 if precision == "fp64" else V32):
        print(f"[variant] batch {precision} v={v}", flush=True)
        monkeypatch.setenv("H1PA_DIFF_VARIANT", str(v))
        keep, xs = [], {}
        bs = M.BatchSolver(prec, iters_per_replay=8, use_graph=False)
        for i, (n, p) in enumerate(specs):
            mesh = M.Mesh.MakeCartesian3D(n, n, n)
            mesh.Perturb(0.15, i)
            fes, a = _form(M, mesh, p, p + 2)
            N = fes.GetNDofs()
            g = torch.Generator(device="cuda").manual_seed(500 + i)
            b = torch.rand(N, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
            ess = fes.GetBoundaryTrueDofs()
            A, X, B = a.FormLinearSystem(ess, torch.zeros(N, dtype=torch.float64,
                                                          device="cuda"), b)
            keep.append((mesh, fes, a, A, B, ess))
            xs[i] = X
            bs.Add(A, B, X, rel_tol=1e-10, max_iter=20000, tag=i)
        assert bs.Run() == 0
        res = {tag: r for tag, r in bs.Poll()}
        out = {i: (xs[i].cpu().numpy(), res[i]) for i in xs}
        assert all(r["converged"] for _, r in out.values())
        if ref is None:
            ref = out
        else:
            for i in out:
                assert out[i][1]["iterations"] == ref[i][1]["iterations"], (v, i)
                assert np.array_equal(out[i][0], ref[i][0]), f"variant {v} problem {i}"
