"""C ABI checks that need no GPU: the library loads, exports every entry point
declared in include/h1pa.h, and its host-side topology (mesh generation,
first-encounter numbering, perturbation) matches the oracle bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "h1pa.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(h1pa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(h1pa):
    lib = ctypes.CDLL(h1pa.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_mirror_binds_every_symbol(h1pa):
    assert sorted(h1pa.SIGNATURES) == declared_symbols()


def test_errors_are_reported(h1pa):
    with pytest.raises(h1pa.H1paError) as e:
        h1pa.Mesh.MakeCartesian3D(0, 1, 1)
    assert e.value.code == h1pa.H1PA_EINVAL
    assert "must be >= 1" in str(e.value)
    with pytest.raises(h1pa.H1paError):
        h1pa.H1_FECollection(9)
    # null handles are refused before any device work (no GPU needed)
    import ctypes as C
    v = C.c_int()
    assert h1pa.lib().h1pa_batch_query(None, C.byref(v)) == h1pa.H1PA_EINVAL
    assert h1pa.lib().h1pa_batch_wait(None, C.byref(v)) == h1pa.H1PA_EINVAL
    assert h1pa.lib().h1pa_mesh_load(None, 0, None) == h1pa.H1PA_EINVAL
    # the device block cache's switch (no device work while nothing is kept)
    with pytest.raises(h1pa.H1paError) as e:
        h1pa.SetDeviceCache(-1)
    assert e.value.code == h1pa.H1PA_EINVAL
    h1pa.SetDeviceCache(1 << 20)
    h1pa.SetDeviceCache(0)
    assert h1pa.DeviceCacheBytes() == 0


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 2, 2), (3, 3, 3), (4, 3, 5), (7, 5, 6),
                                  (16, 16, 16), (9, 2, 3)])
@pytest.mark.parametrize("sfc", [True, False])
def test_topology_matches_oracle(orc, h1pa, dims, sfc):
    om = orc.CartesianMesh(*dims, sfc=sfc)
    pm = h1pa.Mesh.MakeCartesian3D(*dims, sfc_ordering=sfc)
    assert (pm.GetNV(), pm.GetNE(), pm.GetNEdges(), pm.GetNFaces()) == \
        (om.nv, om.ne, om.nedges, om.nfaces)
    t = pm.Topology()
    for k in ("ev", "ee", "eeo", "ef", "efo"):
        assert np.array_equal(getattr(om, k), t[k]), k
    pm.Perturb(0.2, 1234)
    assert np.array_equal(om.vertices(0.2, 1234), pm.GetVertices())


def test_large_topology_is_fast(h1pa):
    """The structured build is O(NE): a 64^3 mesh (262k elements) in well under a second."""
    import time
    t0 = time.time()
    m = h1pa.Mesh.MakeCartesian3D(64, 64, 64)
    assert m.GetNE() == 64 ** 3
    assert time.time() - t0 < 5.0


def test_mixed_precision_argument_errors(h1pa):
    """Argument checks run before any device work, so they hold without a GPU."""
    lib = h1pa.lib()
    o = h1pa.IrOpts(h1pa.PREC_FP32, 1e-10, 1e-4, 10, 100)
    st = h1pa.IrStats()
    assert lib.h1pa_ir_solve(None, None, 0, None, None, ctypes.byref(o), ctypes.byref(st),
                             None) == h1pa.H1PA_EINVAL
    assert b"null argument" in lib.h1pa_last_error()
    assert lib.h1pa_form_mult_f32(None, h1pa.PREC_FP32, None, None, None) == h1pa.H1PA_EINVAL


def test_splitmix64_known_answer():
    """The counter-based generator behind Perturb/Randomize/Fourier seeds is
    SplitMix64: its first output from state 0 is 0xE220A8397B1DCDAF."""
    from test_gpu_serve import splitmix64
    assert splitmix64(0) == 0xE220A8397B1DCDAF
    assert splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


def test_serve_driver_rejects_bad_options():
    import subprocess
    exe = os.path.join(ROOT, "mfem_amd", "h1pa_serve")
    for args in (["--mode", "queue"], ["--pmax", "9"], ["--bogus"], ["--workers", "0"],
                 ["--setup-threads", "9"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=30)
        assert r.returncode == 2


def _grid_arrays(orc, dims, seed):
    """A lexicographic grid with shuffled elements and jittered interior
    vertices, as arrays, with its cube-face boundary quadrilaterals."""
    m = orc.CartesianMesh(*dims, sfc=False)
    xyz = m.vertices(0.15, seed)
    ev = m.ev[np.random.default_rng(seed).permutation(m.ne)].copy()
    faces = [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]
    bq = []
    for e in range(m.ne):
        for f in faces:
            vv = [int(ev[e, k]) for k in f]
            c = xyz[vv]
            if any(np.all(np.abs(c[:, a]) < 1e-12) or np.all(np.abs(c[:, a] - 1) < 1e-12)
                   for a in range(3)):
                bq.append(vv)
    return ev, np.array(bq, np.int32), xyz


@pytest.mark.parametrize("dims,seed", [((1, 1, 1), 1), ((3, 4, 2), 11), ((5, 3, 4), 2)])
def test_arrays_mesh_topology_matches_oracle(orc, h1pa, dims, seed):
    """Mesh from arrays (mesh/mesh.cpp:4891-4930): first-encounter edge and
    face numbering and orientations equal the oracle's on the same arrays;
    the vertices are kept as given."""
    ev, bq, xyz = _grid_arrays(orc, dims, seed)
    om = orc.CartesianMesh(0, 0, 0, elements=(ev, bq, xyz.shape[0]))
    pm = h1pa.Mesh.FromArrays(xyz, ev, bq)
    assert (pm.GetNV(), pm.GetNE(), pm.GetNEdges(), pm.GetNFaces()) == \
        (om.nv, om.ne, om.nedges, om.nfaces)
    t = pm.Topology()
    for k in ("ev", "ee", "eeo", "ef", "efo"):
        assert np.array_equal(getattr(om, k), t[k]), k
    assert np.array_equal(pm.GetVertices(), xyz)


def test_arrays_mesh_argument_errors(h1pa):
    ev = np.arange(8, dtype=np.int32).reshape(1, 8)
    xyz = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                    [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], np.float64)
    none = np.zeros((0, 4), np.int32)
    cases = [(xyz, ev + 1, none, "out of range"),
             (xyz, ev, np.array([[0, 1, 6, 7]], np.int32), "is not an element face"),
             (xyz[[4, 5, 6, 7, 0, 1, 2, 3]].copy(), ev, none, "inverted"),
             (np.where(xyz == 1, np.nan, xyz), ev, none, "finite")]
    for v, e, b, msg in cases:
        with pytest.raises(h1pa.H1paError) as err:
            h1pa.Mesh.FromArrays(v, e, b)
        assert msg in str(err.value)
    m = h1pa.Mesh.FromArrays(xyz, ev, np.array([[0, 3, 2, 1]], np.int32))
    assert (m.GetNE(), m.GetNFaces(), m.GetNEdges()) == (1, 6, 12)
    assert h1pa.lib().h1pa_mesh_set_nodes(None, 2, None, None) == h1pa.H1PA_EINVAL
