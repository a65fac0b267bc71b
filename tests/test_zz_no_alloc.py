"""GPU: h1pa_form_mult_noalloc (the apply whose E-vector lives in the
library's static scratch array) against h1pa_form_mult on the same form,
through the C ABI: within 1e-13 relative (the apply bar), on two calls in a
row, the second on a smaller form, and a form whose E-vector exceeds the
scratch is refused without device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-13


def _form(M, n, p, seed, integ):
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, seed)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    coef = M.SinCosCoefficient()  # one coefficient object: one C form for both integrators
    for it in integ:
        cls = M.DiffusionIntegrator if it == "diffusion" else M.MassIntegrator
        a.AddDomainIntegrator(cls(coef))
    a.Assemble()
    return mesh, fes, a


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_noalloc_matches_mult(h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    lib = M.lib()
    keep = []
    for n, p, integ in ((12, 4, ("diffusion",)), (5, 2, ("diffusion", "mass"))):
        print(f"[case] noalloc n={n} p={p} {integ}", flush=True)
        mesh, fes, a = _form(M, n, p, n, integ)
        keep.append((mesh, fes, a))
        g = torch.Generator(device="cpu").manual_seed(n)
        x = (torch.rand(fes.GetNDofs(), generator=g, dtype=torch.float64) * 2 - 1).cuda()
        y0 = torch.empty_like(x)
        y1 = torch.full_like(x, float("nan"))
        assert lib.h1pa_form_mult(a._single()._h, M._ptr(x), M._ptr(y0), None) == 0
        rc = lib.h1pa_form_mult_noalloc(a._single()._h, M._ptr(x), M._ptr(y1), None)
        assert rc == 0, lib.h1pa_last_error()
        torch.cuda.synchronize()
        assert torch.isfinite(y1).all()
        assert _rel(y1, y0) < RTOL


def test_noalloc_refuses_oversized(h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    lib = M.lib()
    mesh, fes, a = _form(M, 23, 8, 1, ("diffusion",))  # 12167 x 729 E-entries > 8M
    x = torch.zeros(fes.GetNDofs(), dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    rc = lib.h1pa_form_mult_noalloc(a._single()._h, M._ptr(x), M._ptr(y), None)
    assert rc == M.H1PA_EINVAL
    assert b"exceeds the scratch" in lib.h1pa_last_error()
