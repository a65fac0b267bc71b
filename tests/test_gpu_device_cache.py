"""GPU: the library's device block cache (h1pa_set_device_cache; no MFEM
counterpart beyond its pooled device memory types).  With it on, freed blocks
are kept and handed to later allocations: a problem built, applied and
destroyed twice gives bitwise the same results as with the cache off, the
second build takes its blocks from the cache, and turning the cache off frees
them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _apply(M, torch, n, p, seed):
    mesh = M.Mesh.MakeCartesian3D(n, n, n)
    mesh.Perturb(0.2, seed)
    fes = M.FiniteElementSpace(mesh, M.H1_FECollection(p, 3))
    a = M.BilinearForm(fes)
    a.SetAssemblyLevel(M.AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(M.DiffusionIntegrator(M.SinCosCoefficient()))
    a.Assemble()
    N = fes.GetNDofs()
    x = torch.from_numpy(np.random.default_rng(seed).uniform(-1, 1, N)).cuda()
    y = torch.empty_like(x)
    a.Mult(x, y)
    d = torch.empty_like(x)
    a.AssembleDiagonal(d)
    torch.cuda.synchronize()  # no device work pending when the objects go
    return y.cpu().numpy(), d.cpu().numpy()


def test_cache_reuses_blocks_and_changes_nothing(h1pa, torch_cuda):
    M, torch = h1pa, torch_cuda
    print("[cache] n=6 p=3", flush=True)
    ref = _apply(M, torch, 6, 3, 5)
    try:
        M.SetDeviceCache(1 << 30)
        first = _apply(M, torch, 6, 3, 5)
        kept = M.DeviceCacheBytes()
        assert kept > 0, "destroyed objects' blocks are kept"
        second = _apply(M, torch, 6, 3, 5)
        # the same sizes again: the second build reused the kept blocks
        assert M.DeviceCacheBytes() <= kept
        for a_, b_, c_ in zip(ref, first, second):
            assert np.array_equal(a_, b_) and np.array_equal(a_, c_)
        # a lower limit trims what is kept; off frees everything
        M.SetDeviceCache(1 << 16)
        assert M.DeviceCacheBytes() <= (1 << 16)
        _apply(M, torch, 4, 2, 6)
        assert M.DeviceCacheBytes() <= (1 << 16)
    finally:
        M.SetDeviceCache(0)
    assert M.DeviceCacheBytes() == 0
