"""mfem_amd -- MI355X-native H1 partial-assembly operator path.

A thin Python mirror of MFEM's interface for this path (same class and method
names, argument meaning and error behaviour) over the C ABI of libh1pa.so
(include/h1pa.h).  All arithmetic runs in the HIP library; vectors are torch
float64 tensors on the GPU, used only as device memory.

    mesh = Mesh.MakeCartesian3D(16, 16, 16)
    fes = FiniteElementSpace(mesh, H1_FECollection(3, 3))
    a = BilinearForm(fes)
    a.SetAssemblyLevel(AssemblyLevel.PARTIAL)
    a.AddDomainIntegrator(DiffusionIntegrator(ConstantCoefficient(1.0)))
    a.Assemble()
    a.Mult(x, y)

Reference: fem/bilinearform.hpp:31-53 (AssemblyLevel), fem/restriction.hpp:40-108
(ElementRestriction), fem/bilinearform_ext.cpp (PABilinearFormExtension),
linalg/solvers.cpp (CGSolver, OperatorJacobiSmoother).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# The in-tree build; H1PA_LIB names another build of the same library (for
# side-by-side measurements of two versions).
LIB_PATH = os.environ.get("H1PA_LIB") or os.path.join(_HERE, "libh1pa.so")

_lib = None

_i, _i64, _u, _d, _vp = C.c_int, C.c_int64, C.c_uint, C.c_double, C.c_void_p
_pi64 = C.POINTER(C.c_int64)

H1PA_OK, H1PA_EINVAL, H1PA_ENOMEM, H1PA_EHIP, H1PA_ENOTCONV, H1PA_EUNSUP, H1PA_ENONFINITE = \
    range(7)
INTEG_MASS, INTEG_DIFFUSION = 1, 2
COEFF_CONSTANT, COEFF_SINCOS, COEFF_PER_ELEMENT, COEFF_QUADRATURE, COEFF_FOURIER = range(5)
QI_VALUES, QI_DERIVATIVES, QI_DETERMINANTS, QI_PHYSICAL_DERIVATIVES = 1, 2, 4, 8


class H1paError(RuntimeError):
    """A non-zero status from libh1pa (MFEM would call mfem_error here)."""

    def __init__(self, code, msg):
        super().__init__(f"h1pa error {code}: {msg}")
        self.code = code


class FormDesc(C.Structure):
    _fields_ = [("integrators", _u), ("coeff_kind", _i), ("coeff_value", _d),
                ("d_coeff", _vp), ("coeff_seed", C.c_uint64), ("q1d", _i)]


class PcgOpts(C.Structure):
    _fields_ = [("rel_tol", _d), ("abs_tol", _d), ("max_iter", _i), ("check_every", _i)]


class PcgStats(C.Structure):
    _fields_ = [("converged", _i), ("iterations", _i), ("initial_norm", _d),
                ("final_norm", _d), ("seconds", _d)]


class MgOpts(C.Structure):
    _fields_ = [("smoother_order", _i), ("pre_smoothing", _i), ("post_smoothing", _i),
                ("w_cycle", _i), ("coarse_rel_tol", _d), ("coarse_max_iter", _i),
                ("power_iterations", _i), ("power_tolerance", _d), ("power_seed", _i),
                ("cycle_precision", _i), ("coarse_jacobi", _i)]


class IrOpts(C.Structure):
    _fields_ = [("inner_precision", _i), ("rel_tol", _d), ("delta", _d),
                ("max_outer", _i), ("max_inner", _i)]


class IrStats(C.Structure):
    _fields_ = [("converged", _i), ("outer_iterations", _i), ("inner_iterations", _i),
                ("final_norm", _d), ("rel_residual", _d), ("seconds", _d)]


PREC_FP64, PREC_FP32, PREC_BF16_PA = 0, 1, 2


class BatchOpts(C.Structure):
    _fields_ = [("precision", _i), ("iters_per_replay", _i), ("use_graph", _i), ("autotune", _i),
                ("delta", _d), ("max_outer", _i), ("max_inner", _i)]


class BatchResult(C.Structure):
    _fields_ = [("converged", _i), ("iterations", _i), ("outer_iterations", _i),
                ("final_norm", _d), ("rel_residual", _d), ("status", _i)]


# name -> (restype, argtypes); every entry point of include/h1pa.h
SIGNATURES = {
    "h1pa_last_error": (C.c_char_p, []),
    "h1pa_version": (C.c_char_p, []),
    "h1pa_set_device": (_i, [_i]),
    "h1pa_stream_synchronize": (_i, [_vp]),
    "h1pa_mesh_create_cartesian": (_i, [_i, _i, _i, _i, C.POINTER(_vp)]),
    "h1pa_mesh_load": (_i, [C.c_char_p, _i64, C.POINTER(_vp)]),
    "h1pa_mesh_create": (_i, [_i64, _vp, _i64, _vp, _i64, _vp, C.POINTER(_vp)]),
    "h1pa_mesh_set_nodes": (_i, [_vp, _i, _vp, _vp]),
    "h1pa_mesh_destroy": (_i, [_vp]),
    "h1pa_mesh_perturb": (_i, [_vp, _d, C.c_uint64]),
    "h1pa_mesh_set_curvature": (_i, [_vp, _i, _d]),
    "h1pa_mesh_counts": (_i, [_vp, _pi64]),
    "h1pa_mesh_topology": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "h1pa_mesh_vertices": (_i, [_vp, _vp]),
    "h1pa_mesh_geometric_factors": (_i, [_vp, _i, _u, _vp, _vp, _vp, _vp]),
    "h1pa_space_create": (_i, [_vp, _i, C.POINTER(_vp)]),
    "h1pa_space_destroy": (_i, [_vp]),
    "h1pa_space_info": (_i, [_vp, _pi64]),
    "h1pa_space_element_dofs": (_i, [_vp, C.POINTER(_vp)]),
    "h1pa_space_boundary_dofs": (_i, [_vp, C.POINTER(_vp), _pi64]),
    "h1pa_space_restriction": (_i, [_vp, C.POINTER(_vp)]),
    "h1pa_restriction_create": (_i, [_i, _i, _i, _vp, _vp, C.POINTER(_vp)]),
    "h1pa_restriction_destroy": (_i, [_vp]),
    "h1pa_restriction_maps": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "h1pa_restriction_mult": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_restriction_abs_mult": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_restriction_mult_transpose": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_restriction_add_mult_transpose": (_i, [_vp, _vp, _vp, _d, _vp]),
    "h1pa_restriction_abs_mult_transpose": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_restriction_mult_left_inverse": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_restriction_boolean_mask": (_i, [_vp, _vp, _vp]),
    "h1pa_form_create": (_i, [_vp, C.POINTER(FormDesc), C.POINTER(_vp)]),
    "h1pa_form_destroy": (_i, [_vp]),
    "h1pa_form_assemble": (_i, [_vp, _vp]),
    "h1pa_form_mult": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_form_mult_noalloc": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_form_add_mult": (_i, [_vp, _vp, _vp, C.c_double, _vp]),
    "h1pa_form_add_mult_e": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_form_assemble_diagonal": (_i, [_vp, _vp, _vp]),
    "h1pa_form_pa_data": (_i, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "h1pa_form_info": (_i, [_vp, _pi64]),
    "h1pa_form_mult_path": (_i, [_vp, C.POINTER(_i), _vp]),
    "h1pa_form_mult_timed": (_i, [_vp, _vp, _vp, _i, C.POINTER(_d), C.POINTER(_d),
                                  C.POINTER(_d), _vp]),
    "h1pa_form_constrained_mult": (_i, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "h1pa_form_linear_system": (_i, [_vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "h1pa_space_domain_lf": (_i, [_vp, _d, _vp, _vp]),
    "h1pa_space_domain_lf_qf": (_i, [_vp, _d, _vp, _vp, _vp]),
    "h1pa_pcg_solve": (_i, [_vp, _vp, _i64, _vp, _vp, C.POINTER(PcgOpts), C.POINTER(PcgStats),
                            _vp]),
    "h1pa_form_mult_f32": (_i, [_vp, _i, _vp, _vp, _vp]),
    "h1pa_space_qinterp": (_i, [_vp, _i, _i, _u, _i, _vp, _vp, _vp, _vp, _vp]),
    "h1pa_chebyshev_create": (_i, [_vp, _vp, _vp, _i64, _i, _d, _i, _d, _i, C.POINTER(_vp),
                                   _vp]),
    "h1pa_chebyshev_destroy": (_i, [_vp]),
    "h1pa_chebyshev_info": (_i, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    "h1pa_transfer_create": (_i, [_vp, _vp, C.POINTER(_vp)]),
    "h1pa_transfer_destroy": (_i, [_vp]),
    "h1pa_transfer_mult": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_transfer_mult_transpose": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_multigrid_create": (_i, [_vp, _i, C.POINTER(_i), C.POINTER(MgOpts), C.POINTER(_vp),
                                   _vp]),
    "h1pa_multigrid_destroy": (_i, [_vp]),
    "h1pa_multigrid_mult": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_multigrid_info": (_i, [_vp, _pi64, C.POINTER(_d)]),
    "h1pa_pcg_solve_multigrid": (_i, [_vp, _vp, _vp, C.POINTER(PcgOpts), C.POINTER(PcgStats),
                                      _vp]),
    "h1pa_chebyshev_mult": (_i, [_vp, _vp, _vp, _vp]),
    "h1pa_pcg_solve_chebyshev": (_i, [_vp, _vp, _vp, C.POINTER(PcgOpts), C.POINTER(PcgStats),
                                      _vp]),
    "h1pa_batch_create": (_i, [C.POINTER(BatchOpts), C.POINTER(_vp)]),
    "h1pa_batch_destroy": (_i, [_vp]),
    "h1pa_batch_add": (_i, [_vp, _vp, _vp, _i64, _vp, _vp, C.POINTER(PcgOpts), _i64,
                            C.POINTER(_i), _vp]),
    "h1pa_vector_add": (_i, [_vp, C.c_double, _vp, _i64, _vp]),
    "h1pa_vector_randomize": (_i, [_vp, _i64, C.c_uint64, _d, _d, _vp]),
    "h1pa_batch_run": (_i, [_vp, _i, C.POINTER(_i)]),
    "h1pa_batch_launch": (_i, [_vp, C.POINTER(_i)]),
    "h1pa_batch_wait": (_i, [_vp, C.POINTER(_i)]),
    "h1pa_batch_query": (_i, [_vp, C.POINTER(_i)]),
    "h1pa_batch_poll": (_i, [_vp, _i, _pi64, C.POINTER(BatchResult), C.POINTER(_i)]),
    "h1pa_batch_info": (_i, [_vp, _pi64]),
    "h1pa_form_device_bytes": (_i, [_vp, _pi64]),
    "h1pa_set_device_cache": (_i, [C.c_int64]),
    "h1pa_device_cache_bytes": (_i, [_vp]),
    "h1pa_ir_solve": (_i, [_vp, _vp, _i64, _vp, _vp, C.POINTER(IrOpts), C.POINTER(IrStats),
                           _vp]),
}


def lib():
    """Load libh1pa.so (built by mfem_amd.build); fail loudly if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m mfem_amd.build`")
        # torch ships its own HIP runtime; load it first so that libh1pa.so
        # binds to the same libamdhip64 instance (two runtimes in one
        # process make every hipMalloc of the second one fail).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _destroy(fn_name, h):
    """Release a handle; a no-op once the library is gone (interpreter exit)."""
    L = _lib
    if L is None or not h:
        return
    try:
        getattr(L, fn_name)(h)
    except Exception:  # noqa: BLE001 - destructors must not raise
        pass


def _check(status):
    if status != H1PA_OK:
        raise H1paError(status, lib().h1pa_last_error().decode())


def _torch():
    import torch
    return torch


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    if t is None:
        return None
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError("expected a CUDA (HIP) tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _check_vec(t, n, dtype_name="float64"):
    torch = _torch()
    want = getattr(torch, dtype_name)
    if t.dtype != want or t.numel() != n:
        raise ValueError(f"vector must be {dtype_name}[{n}], got {t.dtype}[{t.numel()}]")


def _ess(t):
    """An essential-DOF list as MFEM's Array<int>: an int32 device tensor.
    Integer tensors of other widths are converted (entries must fit in
    int32); anything else is rejected, so an int64 list (torch's default
    index type) is never read as pairs of int32 words."""
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise TypeError("ess_tdof_list must be a CUDA (HIP) integer tensor")
    if t.dtype == torch.int32:
        return t.contiguous()
    if t.dtype in (torch.int64, torch.int16, torch.int8, torch.uint8):
        if t.numel() and (int(t.min()) < -2**31 or int(t.max()) >= 2**31):
            raise ValueError("ess_tdof_list entries do not fit in int32")
        return t.to(torch.int32).contiguous()
    raise TypeError(f"ess_tdof_list must be an integer tensor, got {t.dtype}")


def _wrap_device_int(ptr, n):
    """Copy n int32 values from a device pointer into a new torch tensor."""
    torch = _torch()
    out = torch.empty(int(n), dtype=torch.int32, device="cuda")
    if n:
        from ctypes import c_size_t
        hip = C.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [_vp, _vp, c_size_t, _i]
        hip.hipMemcpy.restype = _i
        rc = hip.hipMemcpy(C.c_void_p(out.data_ptr()), C.c_void_p(ptr), 4 * int(n), 3)
        if rc != 0:
            raise H1paError(H1PA_EHIP, f"hipMemcpy failed ({rc})")
    return out


def _wrap_device_double(ptr, n):
    torch = _torch()
    out = torch.empty(int(n), dtype=torch.float64, device="cuda")
    if n:
        from ctypes import c_size_t
        hip = C.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [_vp, _vp, c_size_t, _i]
        hip.hipMemcpy.restype = _i
        rc = hip.hipMemcpy(C.c_void_p(out.data_ptr()), C.c_void_p(ptr), 8 * int(n), 3)
        if rc != 0:
            raise H1paError(H1PA_EHIP, f"hipMemcpy failed ({rc})")
    return out


# ---------------------------------------------------------------------------
# enums
# ---------------------------------------------------------------------------
class AssemblyLevel:
    """fem/bilinearform.hpp:31-53"""
    LEGACY, FULL, ELEMENT, PARTIAL, NONE = 0, 1, 2, 3, 4


class ElementDofOrdering:
    NATIVE, LEXICOGRAPHIC = 0, 1


class Element:
    HEXAHEDRON = 5


# ---------------------------------------------------------------------------
# mesh
# ---------------------------------------------------------------------------
class Mesh:
    """Hexahedral mesh: Cartesian (Mesh::MakeCartesian3D, mesh/mesh.cpp:4786-4794)
    or read from an MFEM v1.0 mesh file (Mesh(path), mesh/mesh_readers.cpp:40-130)."""

    def __init__(self, handle_or_path, nx=0, ny=0, nz=0):
        if isinstance(handle_or_path, (str, os.PathLike)):
            with open(handle_or_path, "rb") as f:
                text = f.read()
            h = C.c_void_p()
            _check(lib().h1pa_mesh_load(text, len(text), C.byref(h)))
            handle_or_path = h
        self._h = handle_or_path
        self.nx, self.ny, self.nz = nx, ny, nz

    @staticmethod
    def LoadFromFile(path):
        return Mesh(path)

    @staticmethod
    def FromArrays(vertices, elements, boundary):
        """Mesh(vertices, nv, element_indices, CUBE, ..., boundary_indices,
        SQUARE, ...) + FinalizeTopology (mesh/mesh.cpp:4891-4930): host arrays
        vertices [nv][3], elements [ne][8] (MFEM hexahedron vertex order),
        boundary [nbdr][4] quadrilaterals."""
        xyz = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        ev = np.ascontiguousarray(elements, dtype=np.int32).reshape(-1, 8)
        bq = np.ascontiguousarray(boundary, dtype=np.int32).reshape(-1, 4)
        h = C.c_void_p()
        _check(lib().h1pa_mesh_create(xyz.shape[0], xyz.ctypes.data_as(_vp), ev.shape[0],
                                      ev.ctypes.data_as(_vp), bq.shape[0],
                                      bq.ctypes.data_as(_vp) if bq.shape[0] else None,
                                      C.byref(h)))
        return Mesh(h)

    @staticmethod
    def MakeCartesian3D(nx, ny, nz, type=Element.HEXAHEDRON, sx=1.0, sy=1.0, sz=1.0,
                        sfc_ordering=True):
        if type != Element.HEXAHEDRON:
            raise H1paError(H1PA_EUNSUP, "only HEXAHEDRON meshes are supported")
        if (sx, sy, sz) != (1.0, 1.0, 1.0):
            raise H1paError(H1PA_EUNSUP, "only the unit cube is supported")
        h = C.c_void_p()
        _check(lib().h1pa_mesh_create_cartesian(nx, ny, nz, 1 if sfc_ordering else 0,
                                                C.byref(h)))
        return Mesh(h, nx, ny, nz)

    def __del__(self, _d=_destroy):
        _d("h1pa_mesh_destroy", getattr(self, "_h", None))
        self._h = None

    def _counts(self):
        c = (C.c_int64 * 4)()
        _check(lib().h1pa_mesh_counts(self._h, c))
        return list(c)

    def GetNV(self):
        return self._counts()[0]

    def GetNE(self):
        return self._counts()[1]

    def GetNEdges(self):
        return self._counts()[2]

    def GetNFaces(self):
        return self._counts()[3]

    def Perturb(self, amp, seed):
        """Move interior vertices by U(-amp h, amp h) per coordinate (seeded)."""
        _check(lib().h1pa_mesh_perturb(self._h, float(amp), int(seed)))

    def SetCurvature(self, order, amp=0.0):
        """SetCurvature(order) followed by the curving map of amplitude amp."""
        _check(lib().h1pa_mesh_set_curvature(self._h, int(order), float(amp)))

    def SetNodes(self, order, enodes):
        """High-order nodes of an H1_FECollection(order, 3) vector space as
        the lexicographic E-vector [(order+1)^3][3][ne] (FP64, device): what
        the nodal space's ElementRestriction(LEXICOGRAPHIC) gives from
        Mesh::GetNodes() (mesh/mesh.cpp:7200-7240, :15445-15497)."""
        _check_vec(enodes, (order + 1) ** 3 * 3 * self.GetNE())
        _check(lib().h1pa_mesh_set_nodes(self._h, int(order), _ptr(enodes), _stream()))

    def GetVertices(self):
        xyz = np.zeros(3 * self.GetNV())
        _check(lib().h1pa_mesh_vertices(self._h, xyz.ctypes.data_as(_vp)))
        return xyz.reshape(-1, 3)

    def Topology(self):
        ne = self.GetNE()
        ev = np.zeros(8 * ne, np.int32); ee = np.zeros(12 * ne, np.int32)
        eeo = np.zeros(12 * ne, np.int32); ef = np.zeros(6 * ne, np.int32)
        efo = np.zeros(6 * ne, np.int32)
        _check(lib().h1pa_mesh_topology(self._h, *(a.ctypes.data_as(_vp)
                                                   for a in (ev, ee, eeo, ef, efo))))
        return dict(ev=ev.reshape(ne, 8), ee=ee.reshape(ne, 12), eeo=eeo.reshape(ne, 12),
                    ef=ef.reshape(ne, 6), efo=efo.reshape(ne, 6))

    def GetGeometricFactors(self, q1d, flags=QI_VALUES | QI_DERIVATIVES | QI_DETERMINANTS):
        """X[nq][3][ne], J[nq][3][3][ne], detJ[nq][ne] at the q1d^3 GL points."""
        torch = _torch()
        ne, nq = self.GetNE(), q1d ** 3
        X = torch.empty(nq * 3 * ne, dtype=torch.float64, device="cuda") if flags & QI_VALUES else None
        J = torch.empty(nq * 9 * ne, dtype=torch.float64, device="cuda") if flags & QI_DERIVATIVES else None
        dJ = torch.empty(nq * ne, dtype=torch.float64, device="cuda") if flags & QI_DETERMINANTS else None
        _check(lib().h1pa_mesh_geometric_factors(self._h, q1d, flags, _ptr(X), _ptr(J), _ptr(dJ),
                                                 _stream()))
        return X, J, dJ


# ---------------------------------------------------------------------------
# space and restriction
# ---------------------------------------------------------------------------
class H1_FECollection:
    """H1 Gauss-Lobatto collection (fem/fe_coll.cpp:1741)."""

    def __init__(self, p, dim=3):
        if dim != 3:
            raise H1paError(H1PA_EUNSUP, "only dim = 3 is supported")
        if not 1 <= p <= 8:
            raise H1paError(H1PA_EUNSUP, "order must be in [1, 8]")
        self.p, self.dim = p, dim

    def GetOrder(self):
        return self.p


class ElementRestriction:
    """L <-> E maps (fem/restriction.hpp:40-108)."""

    def __init__(self, handle, ne, dof, ndofs, owner=None, owned=True):
        self._h, self.ne, self.dof, self.ndofs = handle, ne, dof, ndofs
        self._owner, self._owned = owner, owned

    @staticmethod
    def FromTable(elem_dof, dof_map, ndofs):
        """ElementRestriction ctor from MFEM's element-to-DOF table + dof_map
        (fem/restriction.cpp:26-114).  Host numpy int32 inputs."""
        ed = np.ascontiguousarray(elem_dof, dtype=np.int32)
        ne, dof = ed.shape
        dm = None if dof_map is None else np.ascontiguousarray(dof_map, dtype=np.int32)
        h = C.c_void_p()
        _check(lib().h1pa_restriction_create(ne, dof, int(ndofs), ed.ctypes.data_as(_vp),
                                             None if dm is None else dm.ctypes.data_as(_vp),
                                             C.byref(h)))
        return ElementRestriction(h, ne, dof, int(ndofs))

    def __del__(self, _d=_destroy):
        if getattr(self, "_owned", False):
            _d("h1pa_restriction_destroy", getattr(self, "_h", None))
        self._h = None

    def Height(self):
        return self.ne * self.dof

    def Width(self):
        return self.ndofs

    def _maps(self):
        g, o, i = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().h1pa_restriction_maps(self._h, C.byref(g), C.byref(o), C.byref(i)))
        return g.value, o.value, i.value

    def GatherMap(self):
        return _wrap_device_int(self._maps()[0], self.ne * self.dof)

    def Offsets(self):
        return _wrap_device_int(self._maps()[1], self.ndofs + 1)

    def Indices(self):
        return _wrap_device_int(self._maps()[2], self.ne * self.dof)

    def Mult(self, x, y):
        _check_vec(x, self.ndofs); _check_vec(y, self.ne * self.dof)
        _check(lib().h1pa_restriction_mult(self._h, _ptr(x), _ptr(y), _stream()))

    def AbsMult(self, x, y):
        _check_vec(x, self.ndofs); _check_vec(y, self.ne * self.dof)
        _check(lib().h1pa_restriction_abs_mult(self._h, _ptr(x), _ptr(y), _stream()))

    def MultTranspose(self, x, y):
        _check_vec(x, self.ne * self.dof); _check_vec(y, self.ndofs)
        _check(lib().h1pa_restriction_mult_transpose(self._h, _ptr(x), _ptr(y), _stream()))

    def AddMultTranspose(self, x, y, a=1.0):
        _check_vec(x, self.ne * self.dof); _check_vec(y, self.ndofs)
        _check(lib().h1pa_restriction_add_mult_transpose(self._h, _ptr(x), _ptr(y), float(a),
                                                         _stream()))

    def AbsMultTranspose(self, x, y):
        _check_vec(x, self.ne * self.dof); _check_vec(y, self.ndofs)
        _check(lib().h1pa_restriction_abs_mult_transpose(self._h, _ptr(x), _ptr(y), _stream()))

    def MultLeftInverse(self, x, y):
        _check_vec(x, self.ne * self.dof); _check_vec(y, self.ndofs)
        _check(lib().h1pa_restriction_mult_left_inverse(self._h, _ptr(x), _ptr(y), _stream()))

    def BooleanMask(self, y):
        """y (E-vector) = 1 at each L-DOF's first E-entry, 0 elsewhere
        (fem/restriction.cpp:255-290)."""
        _check_vec(y, self.ne * self.dof)
        _check(lib().h1pa_restriction_boolean_mask(self._h, _ptr(y), _stream()))


class FiniteElementSpace:
    """FiniteElementSpace(mesh, H1_FECollection) on a conforming hex mesh."""

    def __init__(self, mesh, fec):
        self.mesh, self.fec = mesh, fec
        h = C.c_void_p()
        _check(lib().h1pa_space_create(mesh._h, fec.p, C.byref(h)))
        self._h = h
        info = (C.c_int64 * 4)()
        _check(lib().h1pa_space_info(h, info))
        self.p, self.ne, self.dof, self.ndofs = (int(v) for v in info)
        r = C.c_void_p()
        _check(lib().h1pa_space_restriction(h, C.byref(r)))
        self._restr = ElementRestriction(r, self.ne, self.dof, self.ndofs, owner=self, owned=False)
        self._restr_native = None

    def __del__(self, _d=_destroy):
        self._restr = None
        self._restr_native = None
        _d("h1pa_space_destroy", getattr(self, "_h", None))
        self._h = None

    def GetNE(self):
        return self.ne

    def GetNDofs(self):
        return self.ndofs

    def GetVSize(self):
        return self.ndofs

    def GetTrueVSize(self):
        return self.ndofs

    def GetElementRestriction(self, ordering=ElementDofOrdering.LEXICOGRAPHIC):
        """ElementRestriction in LEXICOGRAPHIC order (the space's own, used by
        PA) or NATIVE order (built once from the element-to-DOF table without
        the dof_map reordering, fem/restriction.cpp:60-98)."""
        if ordering == ElementDofOrdering.LEXICOGRAPHIC:
            return self._restr
        if ordering != ElementDofOrdering.NATIVE:
            raise H1paError(H1PA_EINVAL, "ordering must be NATIVE or LEXICOGRAPHIC")
        if self._restr_native is None:
            ed = self.GetElementToDofTable().cpu().numpy()
            self._restr_native = ElementRestriction.FromTable(ed, None, self.ndofs)
        return self._restr_native

    def GetElementToDofTable(self):
        """Native-order element-to-DOF table as an int32 [ne, dof] device tensor."""
        d = C.c_void_p()
        _check(lib().h1pa_space_element_dofs(self._h, C.byref(d)))
        return _wrap_device_int(d.value, self.ne * self.dof).view(self.ne, self.dof)

    def GetBoundaryTrueDofs(self):
        d, n = C.c_void_p(), C.c_int64()
        _check(lib().h1pa_space_boundary_dofs(self._h, C.byref(d), C.byref(n)))
        return _wrap_device_int(d.value, n.value)

    def _boundary_raw(self):
        d, n = C.c_void_p(), C.c_int64()
        _check(lib().h1pa_space_boundary_dofs(self._h, C.byref(d), C.byref(n)))
        return d, n.value

    def GetQuadratureInterpolator(self, q1d, vdim=1):
        """QuadratureInterpolator at the q1d^3 Gauss-Legendre points
        (FiniteElementSpace::GetQuadratureInterpolator(ir) with the tensor
        rule of q1d points per direction); vdim: components of the E-vectors
        it will interpolate."""
        return QuadratureInterpolator(self, q1d, vdim)


class QVectorLayout:
    byNODES, byVDIM = 0, 1


class QuadratureInterpolator:
    """QuadratureInterpolator (fem/quadinterpolator.hpp:21-160) for
    lexicographic E-vectors [D^3][vdim][ne] of an H1 space."""

    VALUES, DERIVATIVES, DETERMINANTS, PHYSICAL_DERIVATIVES = 1, 2, 4, 8

    def __init__(self, fes, q1d, vdim=1):
        self.fes, self.q1d, self.vdim = fes, int(q1d), int(vdim)
        self.layout = QVectorLayout.byNODES

    def SetOutputLayout(self, layout):
        self.layout = int(layout)

    def GetOutputLayout(self):
        return self.layout

    def Mult(self, e_vec, eval_flags, q_val, q_der, q_det):
        fes, nq = self.fes, self.q1d ** 3
        _check_vec(e_vec, fes.dof * self.vdim * fes.ne)
        if eval_flags & self.VALUES:
            _check_vec(q_val, nq * self.vdim * fes.ne)
        if eval_flags & (self.DERIVATIVES | self.PHYSICAL_DERIVATIVES):
            _check_vec(q_der, nq * self.vdim * 3 * fes.ne)
        if eval_flags & self.DETERMINANTS:
            _check_vec(q_det, nq * fes.ne)
        opt = lambda t: _ptr(t) if t is not None else None  # noqa: E731
        _check(lib().h1pa_space_qinterp(fes._h, self.q1d, self.vdim, int(eval_flags),
                                        self.layout, _ptr(e_vec), opt(q_val), opt(q_der),
                                        opt(q_det), _stream()))

    def Values(self, e_vec, q_val):
        self.Mult(e_vec, self.VALUES, q_val, None, None)

    def Derivatives(self, e_vec, q_der):
        self.Mult(e_vec, self.DERIVATIVES, None, q_der, None)

    def PhysDerivatives(self, e_vec, q_der):
        self.Mult(e_vec, self.PHYSICAL_DERIVATIVES, None, q_der, None)

    def Determinants(self, e_vec, q_det):
        self.Mult(e_vec, self.DETERMINANTS, None, None, q_det)


# ---------------------------------------------------------------------------
# coefficients and integrators
# ---------------------------------------------------------------------------
class Coefficient:
    kind = COEFF_CONSTANT
    value = 1.0
    data = None
    seed = 0


class ConstantCoefficient(Coefficient):
    def __init__(self, value=1.0):
        self.kind, self.value = COEFF_CONSTANT, float(value)


class SinCosCoefficient(Coefficient):
    """k(x) = 1 + 0.5 sin(2 pi x) cos(2 pi y) sin(2 pi z), evaluated at the
    physical quadrature points (a FunctionCoefficient in MFEM terms)."""

    def __init__(self):
        self.kind = COEFF_SINCOS


class PWConstCoefficient(Coefficient):
    """One value per element (device float64 tensor of size NE)."""

    def __init__(self, values):
        self.kind, self.data = COEFF_PER_ELEMENT, values


class QuadratureFunctionCoefficient(Coefficient):
    """Values at every quadrature point, [q1d^3][ne] (device float64)."""

    def __init__(self, values):
        self.kind, self.data = COEFF_QUADRATURE, values


class FourierCoefficient(Coefficient):
    """k = exp(g), g = sum of 8 seeded Fourier modes (amplitude 0.5/sqrt 8 each)."""

    def __init__(self, seed):
        self.kind, self.seed = COEFF_FOURIER, int(seed)


class BilinearFormIntegrator:
    flag = 0

    def __init__(self, coeff=None, q1d=0):
        self.coeff = coeff if coeff is not None else ConstantCoefficient(1.0)
        self.q1d = q1d


class DiffusionIntegrator(BilinearFormIntegrator):
    flag = INTEG_DIFFUSION


class MassIntegrator(BilinearFormIntegrator):
    flag = INTEG_MASS


# ---------------------------------------------------------------------------
# forms, operators, solvers
# ---------------------------------------------------------------------------
class _PAForm:
    """One C form (a set of integrators sharing a coefficient and q1d)."""

    def __init__(self, fes, flags, coeff, q1d):
        # the device reads the coefficient data at fixed sizes: check them here
        if coeff.kind in (COEFF_PER_ELEMENT, COEFF_QUADRATURE):
            q = q1d if q1d > 0 else fes.p + 2
            want = fes.ne if coeff.kind == COEFF_PER_ELEMENT else q ** 3 * fes.ne
            what = ("PWConstCoefficient" if coeff.kind == COEFF_PER_ELEMENT
                    else "QuadratureFunctionCoefficient")
            _ptr(coeff.data)
            try:
                _check_vec(coeff.data, want)
            except ValueError as e:
                raise ValueError(f"{what}: {e}") from None
        desc = FormDesc()
        desc.integrators = flags
        desc.coeff_kind = coeff.kind
        desc.coeff_value = coeff.value
        desc.d_coeff = _ptr(coeff.data).value if coeff.data is not None else None
        desc.coeff_seed = coeff.seed
        desc.q1d = q1d
        self._keep = coeff.data
        h = C.c_void_p()
        _check(lib().h1pa_form_create(fes._h, C.byref(desc), C.byref(h)))
        self._h = h

    def __del__(self, _d=_destroy):
        _d("h1pa_form_destroy", getattr(self, "_h", None))
        self._h = None


class BilinearForm:
    """BilinearForm with AssemblyLevel::PARTIAL (fem/bilinearform.cpp:109-136)."""

    def __init__(self, fes):
        self.fes = fes
        self._level = None
        self._integs = []
        self._forms = []
        self._tmp = None

    def SetAssemblyLevel(self, level):
        if self._level is not None:
            raise H1paError(H1PA_EINVAL, "AssemblyLevel can only be set once")
        if level != AssemblyLevel.PARTIAL:
            raise H1paError(H1PA_EUNSUP, "only AssemblyLevel.PARTIAL is provided")
        self._level = level

    def AddDomainIntegrator(self, integ):
        self._integs.append(integ)

    def Assemble(self):
        if self._level != AssemblyLevel.PARTIAL:
            raise H1paError(H1PA_EINVAL, "call SetAssemblyLevel(PARTIAL) first")
        if not self._integs:
            raise H1paError(H1PA_EINVAL, "no integrators")
        # integrators with the same coefficient object and q1d share one C form
        groups = {}
        for it in self._integs:
            key = (id(it.coeff), it.q1d)
            g = groups.setdefault(key, [0, it.coeff, it.q1d])
            if g[0] & it.flag:
                raise H1paError(H1PA_EUNSUP, "duplicate integrator with the same coefficient")
            g[0] |= it.flag
        key = [(f, id(c), q) for f, c, q in groups.values()]
        # re-assembly keeps the C forms (as MFEM re-assembles in place); only
        # a changed integrator set makes new ones
        if not self._forms or getattr(self, "_form_key", None) != key:
            self._forms = [_PAForm(self.fes, f, c, q) for f, c, q in groups.values()]
            self._form_key = key
        for f in self._forms:
            _check(lib().h1pa_form_assemble(f._h, _stream()))

    def Height(self):
        return self.fes.ndofs

    def Width(self):
        return self.fes.ndofs

    def _scratch(self):
        torch = _torch()
        if self._tmp is None:
            self._tmp = torch.empty(self.fes.ndofs, dtype=torch.float64, device="cuda")
        return self._tmp

    def Mult(self, x, y):
        n = self.fes.ndofs
        _check_vec(x, n); _check_vec(y, n)
        _check(lib().h1pa_form_mult(self._forms[0]._h, _ptr(x), _ptr(y), _stream()))
        for f in self._forms[1:]:
            _check(lib().h1pa_form_add_mult(f._h, _ptr(x), _ptr(y), 1.0, _stream()))

    def AddMult(self, x, y, a=1.0):
        """y += a A x (Operator::AddMult, linalg/operator.cpp:51-56)."""
        n = self.fes.ndofs
        _check_vec(x, n); _check_vec(y, n)
        for f in self._forms:
            _check(lib().h1pa_form_add_mult(f._h, _ptr(x), _ptr(y), float(a), _stream()))

    def MultTranspose(self, x, y):
        """y = A^T x, which is A x: the PA mass and diffusion operators are
        symmetric (fem/integ/bilininteg_diffusion_pa.cpp:95-106)."""
        self.Mult(x, y)

    def AddMultTranspose(self, x, y, a=1.0):
        """y += a A^T x = y + a A x."""
        self.AddMult(x, y, a)

    def MultTimed(self, x, y, steps):
        """`steps` applies with HIP events around every launch; returns
        (total_ms, element_kernel_ms_avg, scatter_ms_avg)."""
        f = self._single()
        tot, te, ts = C.c_double(), C.c_double(), C.c_double()
        _check(lib().h1pa_form_mult_timed(f._h, _ptr(x), _ptr(y), int(steps), C.byref(tot),
                                          C.byref(te), C.byref(ts), _stream()))
        return tot.value, te.value, ts.value

    def MultLowPrecision(self, x, y, precision=PREC_FP32):
        """y = A x with FP32 vectors and FP32 (PREC_FP32) or BF16-stored
        (PREC_BF16_PA) PA data; no MFEM counterpart."""
        n = self.fes.ndofs
        _check_vec(x, n, "float32"); _check_vec(y, n, "float32")
        _check(lib().h1pa_form_mult_f32(self._single()._h, int(precision), _ptr(x), _ptr(y),
                                        _stream()))

    def AddMultPA_E(self, xE, yE):
        """yE += A_E xE on E-vectors (BilinearFormIntegrator::AddMultPA)."""
        for f in self._forms:
            _check(lib().h1pa_form_add_mult_e(f._h, _ptr(xE), _ptr(yE), _stream()))

    def AssembleDiagonal(self, diag):
        _check_vec(diag, self.fes.ndofs)
        _check(lib().h1pa_form_assemble_diagonal(self._forms[0]._h, _ptr(diag), _stream()))
        for f in self._forms[1:]:
            t = self._scratch()
            _check(lib().h1pa_form_assemble_diagonal(f._h, _ptr(t), _stream()))
            _check(lib().h1pa_vector_add(_ptr(diag), 1.0, _ptr(t), t.numel(), _stream()))

    def PAData(self):
        """(diffusion D[q^3*6*ne], mass v[q^3*ne]) copies of the first form's PA data."""
        dd, dm = C.c_void_p(), C.c_void_p()
        f = self._forms[0]
        _check(lib().h1pa_form_pa_data(f._h, C.byref(dd), C.byref(dm)))
        info = (C.c_int64 * 4)()
        _check(lib().h1pa_form_info(f._h, info))
        q = int(info[0])
        nq = q ** 3 * self.fes.ne
        return (_wrap_device_double(dd.value, 6 * nq) if dd.value else None,
                _wrap_device_double(dm.value, nq) if dm.value else None)

    def Info(self):
        """q1d, integrator bitmask, ndofs and the selected diffusion-apply
        variant (-1: none) of the first form."""
        info = (C.c_int64 * 4)()
        _check(lib().h1pa_form_info(self._forms[0]._h, info))
        return dict(q1d=info[0], integrators=info[1], ndofs=info[2], variant=info[3])

    def MultPath(self):
        """'columns' when Mult runs the fused column apply, 'element+scatter'
        when it runs the element kernels and the E->L scatter."""
        v = C.c_int()
        _check(lib().h1pa_form_mult_path(self._single()._h, C.byref(v), _stream()))
        return "columns" if v.value else "element+scatter"

    def DeviceBytes(self):
        """Device bytes held by the form(s), the space and the mesh."""
        tot = 0
        for f in self._forms:
            v = C.c_int64()
            _check(lib().h1pa_form_device_bytes(f._h, C.byref(v)))
            tot += v.value
        return tot

    def _single(self):
        if len(self._forms) != 1:
            raise H1paError(H1PA_EUNSUP, "this call needs all integrators to share a coefficient")
        return self._forms[0]

    def FormLinearSystem(self, ess_tdof_list, x, b):
        """Returns (A, X, B) as MFEM's FormLinearSystem (copy_interior = 0)."""
        n = self.fes.ndofs
        torch = _torch()
        X = torch.empty(n, dtype=torch.float64, device="cuda")
        B = torch.empty(n, dtype=torch.float64, device="cuda")
        f = self._single()
        ess_tdof_list = _ess(ess_tdof_list)
        _check(lib().h1pa_form_linear_system(f._h, _ptr(ess_tdof_list), ess_tdof_list.numel(),
                                             _ptr(x), _ptr(b), _ptr(X), _ptr(B), _stream()))
        return ConstrainedOperator(self, ess_tdof_list), X, B

    def RecoverFEMSolution(self, X, b, x):
        x.copy_(X)


class ConstrainedOperator:
    """DIAG_ONE constrained operator (linalg/operator.cpp:526-662)."""

    def __init__(self, form, ess):
        self.form, self.ess = form, _ess(ess)

    def Height(self):
        return self.form.fes.ndofs

    def Mult(self, x, y):
        f = self.form._single()
        _check(lib().h1pa_form_constrained_mult(f._h, _ptr(self.ess), self.ess.numel(), _ptr(x),
                                                _ptr(y), _stream()))


class OperatorJacobiSmoother:
    """Jacobi from the PA diagonal, ess rows 1 (linalg/solvers.cpp:324-453)."""

    def __init__(self, a, ess_tdof_list, damping=1.0):
        if damping != 1.0:
            raise H1paError(H1PA_EUNSUP, "only damping = 1 is provided")
        self.a, self.ess = a, _ess(ess_tdof_list)


class OperatorChebyshevSmoother:
    """OperatorChebyshevSmoother (linalg/solvers.hpp:502-585,
    solvers.cpp:455-656) on the DIAG_ONE operator from FormLinearSystem.

    OperatorChebyshevSmoother(A, diag, ess, order, power_iterations=10,
    power_tolerance=1e-8, power_seed=12345) estimates the largest eigenvalue
    of diag^-1 A with MFEM's power method from Vector::Randomize(power_seed);
    max_eig_estimate=... uses the given bound instead (MFEM's other
    constructor).  diag=None takes the form's AssembleDiagonal."""

    def __init__(self, A, diag, ess_tdof_list, order, power_iterations=10,
                 power_tolerance=1e-8, power_seed=12345, max_eig_estimate=None):
        if not isinstance(A, ConstrainedOperator):
            raise H1paError(H1PA_EUNSUP, "the operator must come from FormLinearSystem")
        f = A.form._single()
        n = A.form.fes.ndofs
        if diag is not None:
            _check_vec(diag, n)
        ess_tdof_list = _ess(ess_tdof_list)
        self.A, self.ess, self.order = A, ess_tdof_list, int(order)
        self._form = f  # keeps the form alive
        h = _vp()
        _check(lib().h1pa_chebyshev_create(
            f._h, _ptr(diag) if diag is not None else None, _ptr(ess_tdof_list),
            ess_tdof_list.numel(), self.order,
            float(max_eig_estimate) if max_eig_estimate else 0.0, int(power_iterations),
            float(power_tolerance), int(power_seed), C.byref(h), _stream()))
        self._h = h

    def __del__(self, _d=_destroy):
        _d("h1pa_chebyshev_destroy", getattr(self, "_h", None))
        self._h = None

    def GetMaxEigEstimate(self):
        e = _d()
        _check(lib().h1pa_chebyshev_info(self._h, C.byref(e), None))
        return e.value

    def GetCoefficients(self):
        c = (_d * self.order)()
        _check(lib().h1pa_chebyshev_info(self._h, None, c))
        return list(c)

    def Mult(self, x, y):
        n = self.A.form.fes.ndofs
        _check_vec(x, n)
        _check_vec(y, n)
        _check(lib().h1pa_chebyshev_mult(self._h, _ptr(x), _ptr(y), _stream()))


class TransferOperator:
    """TransferOperator(lFESpace, hFESpace) between two H1 spaces of one mesh:
    MFEM picks TensorProductPRefinementTransferOperator for them
    (fem/transfer.cpp:2619-2990).  Mult: low -> high (prolongation),
    MultTranspose: high -> low (restriction)."""

    def __init__(self, lfes, hfes):
        h = _vp()
        _check(lib().h1pa_transfer_create(lfes._h, hfes._h, C.byref(h)))
        self._h, self.lfes, self.hfes = h, lfes, hfes

    def __del__(self, _d=_destroy):
        _d("h1pa_transfer_destroy", getattr(self, "_h", None))
        self._h = None

    def Height(self):
        return self.hfes.ndofs

    def Width(self):
        return self.lfes.ndofs

    def Mult(self, x, y):
        _check_vec(x, self.lfes.ndofs); _check_vec(y, self.hfes.ndofs)
        _check(lib().h1pa_transfer_mult(self._h, _ptr(x), _ptr(y), _stream()))

    def MultTranspose(self, x, y):
        _check_vec(x, self.hfes.ndofs); _check_vec(y, self.lfes.ndofs)
        _check(lib().h1pa_transfer_mult_transpose(self._h, _ptr(x), _ptr(y), _stream()))


class PMultigrid:
    """Multigrid over order-refined levels of the form's mesh: MFEM's
    GeometricMultigrid on a FiniteElementSpaceHierarchy built with
    AddOrderRefinedLevel, with the levels of ex26's DiffusionMultigrid
    (examples/ex26.cpp:42-110): PA forms with the fine form's integrators and
    coefficient, OperatorChebyshevSmoother(order) on levels >= 1, CGSolver
    without a preconditioner (rel_tol coarse_rel_tol, max coarse_max_iter) on
    the coarsest, every level constrained on its boundary DOFs.
    orders: increasing, ending with the fine order (ex26: 1, 2, 4, ...).
    Mult(x, y): one cycle (MultigridBase::Mult, fem/multigrid.cpp:106-145).
    cycle_precision=PREC_BF16_PA: the cycle's applies read BF16 copies of the
    levels' PA data (FP64 vectors): a mixed-precision preconditioner.
    coarse_jacobi=True: Jacobi-preconditioned coarse CG (ex26: none)."""

    def __init__(self, a, orders, smoother_order=2, pre_smoothing=1, post_smoothing=1,
                 cycle="V", coarse_rel_tol=1e-2, coarse_max_iter=200, power_iterations=10,
                 power_tolerance=1e-8, power_seed=12345, cycle_precision=PREC_FP64,
                 coarse_jacobi=False):
        f = a._single()
        orders = [int(o) for o in orders]
        arr = (_i * len(orders))(*orders)
        o = MgOpts(int(smoother_order), int(pre_smoothing), int(post_smoothing),
                   1 if cycle.upper().startswith("W") else 0, float(coarse_rel_tol),
                   int(coarse_max_iter), int(power_iterations), float(power_tolerance),
                   int(power_seed), int(cycle_precision), 1 if coarse_jacobi else 0)
        h = _vp()
        _check(lib().h1pa_multigrid_create(f._h, len(orders), arr, C.byref(o), C.byref(h),
                                           _stream()))
        self._h, self.form, self.orders = h, a, orders

    def __del__(self, _d=_destroy):
        _d("h1pa_multigrid_destroy", getattr(self, "_h", None))
        self._h = None

    def NumLevels(self):
        return len(self.orders)

    def Mult(self, x, y):
        n = self.form.fes.ndofs
        _check_vec(x, n); _check_vec(y, n)
        _check(lib().h1pa_multigrid_mult(self._h, _ptr(x), _ptr(y), _stream()))

    def Info(self):
        info = (C.c_int64 * 4)()
        eig = (_d * len(self.orders))()
        _check(lib().h1pa_multigrid_info(self._h, info, eig))
        return dict(levels=info[0], cycles=info[1], coarse_solves=info[2],
                    coarse_iterations=info[3], max_eig=list(eig))


class CGSolver:
    """CGSolver (linalg/solvers.hpp:626, solvers.cpp:869-1050) with the PA
    operator and OperatorJacobiSmoother or OperatorChebyshevSmoother, run
    fully on the device."""

    def __init__(self):
        self.rel_tol, self.abs_tol, self.max_iter = 0.0, 0.0, 10
        self.op = None
        self.prec = None
        self.stats = PcgStats()
        self.check_every = 8

    def SetRelTol(self, t):
        self.rel_tol = float(t)

    def SetAbsTol(self, t):
        self.abs_tol = float(t)

    def SetMaxIter(self, n):
        self.max_iter = int(n)

    def SetPrintLevel(self, _lvl):
        pass

    def SetOperator(self, op):
        if not isinstance(op, ConstrainedOperator):
            raise H1paError(H1PA_EUNSUP, "the operator must come from FormLinearSystem")
        self.op = op

    def SetPreconditioner(self, prec):
        if not isinstance(prec, (OperatorJacobiSmoother, OperatorChebyshevSmoother, PMultigrid)):
            raise H1paError(H1PA_EUNSUP, "OperatorJacobiSmoother, OperatorChebyshevSmoother "
                                         "and PMultigrid are provided")
        self.prec = prec

    def Mult(self, b, x):
        if self.op is None or self.prec is None:
            raise H1paError(H1PA_EINVAL, "SetOperator and SetPreconditioner first")
        f = self.op.form._single()
        o = PcgOpts(self.rel_tol, self.abs_tol, self.max_iter, self.check_every)
        st = self.stats
        if isinstance(self.prec, PMultigrid):
            if self.prec.form is not self.op.form:
                raise H1paError(H1PA_EUNSUP, "the multigrid must be built on this operator's form")
            ess = self.op.form.fes.GetBoundaryTrueDofs()
            if self.op.ess.numel() != ess.numel() or not bool((self.op.ess == ess).all()):
                raise H1paError(H1PA_EUNSUP, "the multigrid constrains the boundary DOFs: "
                                             "the operator's list must be GetBoundaryTrueDofs()")
            rc = lib().h1pa_pcg_solve_multigrid(self.prec._h, _ptr(b), _ptr(x), C.byref(o),
                                                C.byref(st), _stream())
        elif isinstance(self.prec, OperatorChebyshevSmoother):
            if self.prec.A.form is not self.op.form:
                raise H1paError(H1PA_EUNSUP, "the smoother must be built on this operator")
            rc = lib().h1pa_pcg_solve_chebyshev(self.prec._h, _ptr(b), _ptr(x), C.byref(o),
                                                C.byref(st), _stream())
        else:
            rc = lib().h1pa_pcg_solve(f._h, _ptr(self.op.ess), self.op.ess.numel(), _ptr(b),
                                      _ptr(x), C.byref(o), C.byref(st), _stream())
        if rc not in (H1PA_OK, H1PA_ENOTCONV):
            _check(rc)

    def GetNumIterations(self):
        return self.stats.iterations

    def GetFinalNorm(self):
        return self.stats.final_norm

    def GetConverged(self):
        return bool(self.stats.converged)


class IRSolver:
    """Mixed-precision solve (no MFEM counterpart): inner Jacobi-CG with FP32
    vectors and FP32 or BF16-stored PA data, FP64 iterative refinement by
    reliable updates (FP64 residual and solution update, search direction
    kept).  Stops on CGSolver's measure (B r, r) <= rel_tol^2 (B b, b) taken
    on the FP64 true residual.  Same Set*/Mult shape as CGSolver."""

    def __init__(self, inner_precision=PREC_FP32):
        self.inner_precision = int(inner_precision)
        self.rel_tol, self.delta = 1e-10, 0.1
        self.max_outer, self.max_inner = 200, 100000
        self.op = None
        self.stats = IrStats()

    def SetRelTol(self, t):
        self.rel_tol = float(t)

    def SetDelta(self, t):
        """Refinement trigger: a step whenever the inner residual fell by t."""
        self.delta = float(t)

    def SetMaxIter(self, outer, inner=None):
        self.max_outer = int(outer)
        if inner is not None:
            self.max_inner = int(inner)

    def SetOperator(self, op):
        if not isinstance(op, ConstrainedOperator):
            raise H1paError(H1PA_EUNSUP, "the operator must come from FormLinearSystem")
        self.op = op

    def SetPreconditioner(self, prec):
        if not isinstance(prec, OperatorJacobiSmoother):
            raise H1paError(H1PA_EUNSUP, "only OperatorJacobiSmoother is provided")

    def Mult(self, b, x):
        if self.op is None:
            raise H1paError(H1PA_EINVAL, "SetOperator first")
        f = self.op.form._single()
        o = IrOpts(self.inner_precision, self.rel_tol, self.delta, self.max_outer,
                   self.max_inner)
        rc = lib().h1pa_ir_solve(f._h, _ptr(self.op.ess), self.op.ess.numel(), _ptr(b), _ptr(x),
                                 C.byref(o), C.byref(self.stats), _stream())
        if rc not in (H1PA_OK, H1PA_ENOTCONV):
            _check(rc)

    def GetConverged(self):
        return bool(self.stats.converged)


def Randomize(x, seed, lo=0.0, hi=1.0):
    """Vector::Randomize on the device (linalg/vector.cpp:955-966), with a
    counter-based generator: x[i] = lo + (hi - lo) u_i, reproducible."""
    _check_vec(x, x.numel())
    _check(lib().h1pa_vector_randomize(_ptr(x), x.numel(), int(seed) & ((1 << 64) - 1),
                                       float(lo), float(hi), _stream()))


def SetDeviceCache(max_bytes):
    """Keep up to max_bytes of freed device blocks per device for reuse
    instead of hipFree (which synchronizes the device); 0 turns it off and
    frees them.  With it on, destroy no object while device work on its
    buffers is pending."""
    _check(lib().h1pa_set_device_cache(int(max_bytes)))


def DeviceCacheBytes():
    v = C.c_int64()
    _check(lib().h1pa_device_cache_bytes(C.byref(v)))
    return v.value


class BatchSolver:
    """Many independent Jacobi-PCG (or mixed-precision) solves advanced
    together on the device; no MFEM counterpart.  Each added solve keeps the
    CGSolver settings it was added with; Run advances all running solves by
    `iters_per_replay` iterations per replay of a captured HIP graph; Poll
    returns and retires the finished ones.  The operator's form and the B, X
    tensors are kept alive here until the solve is polled."""

    def __init__(self, precision=PREC_FP64, iters_per_replay=8, use_graph=True, autotune=True,
                 delta=0.1, max_outer=200, max_inner=100000):
        o = BatchOpts(int(precision), int(iters_per_replay), int(bool(use_graph)),
                      int(bool(autotune)), float(delta), int(max_outer), int(max_inner))
        h = C.c_void_p()
        _check(lib().h1pa_batch_create(C.byref(o), C.byref(h)))
        self._h = h
        self._keep = {}
        self._next_tag = 0

    def __del__(self, _d=_destroy):
        _d("h1pa_batch_destroy", getattr(self, "_h", None))
        self._h = None

    def Add(self, A, B, X, rel_tol=1e-12, abs_tol=0.0, max_iter=10000, tag=None):
        """Add the solve A X = B (A from FormLinearSystem); returns its tag."""
        if not isinstance(A, ConstrainedOperator):
            raise H1paError(H1PA_EUNSUP, "the operator must come from FormLinearSystem")
        n = A.Height()
        _check_vec(B, n); _check_vec(X, n)
        if tag is None:
            tag = self._next_tag
        self._next_tag = max(self._next_tag, int(tag)) + 1
        f = A.form._single()
        o = PcgOpts(float(rel_tol), float(abs_tol), int(max_iter), 8)
        slot = C.c_int()
        _check(lib().h1pa_batch_add(self._h, f._h, _ptr(A.ess), A.ess.numel(), _ptr(B), _ptr(X),
                                    C.byref(o), int(tag), C.byref(slot), _stream()))
        self._keep[int(tag)] = (A, B, X, f)
        return int(tag)

    def Run(self, max_replays=1 << 30):
        """Advance until nothing runs or max_replays replays; returns the
        number of solves still running."""
        n = C.c_int()
        _check(lib().h1pa_batch_run(self._h, int(max_replays), C.byref(n)))
        return n.value

    def Launch(self):
        """Start one replay without waiting; False when nothing runs."""
        v = C.c_int()
        _check(lib().h1pa_batch_launch(self._h, C.byref(v)))
        return bool(v.value)

    def Idle(self):
        """True when no replay is in flight or it has finished (never blocks)."""
        v = C.c_int()
        _check(lib().h1pa_batch_query(self._h, C.byref(v)))
        return bool(v.value)

    def Wait(self):
        """Wait for the replay in flight; returns the number still running."""
        n = C.c_int()
        _check(lib().h1pa_batch_wait(self._h, C.byref(n)))
        return n.value

    def Poll(self, cap=4096):
        """[(tag, {converged, iterations, outer_iterations, final_norm,
        rel_residual, status})] for the solves that finished since the last
        call; status is H1PA_OK, H1PA_ENOTCONV or H1PA_ENONFINITE (a
        non-finite product retired that solve alone)."""
        tags = (C.c_int64 * cap)()
        res = (BatchResult * cap)()
        n = C.c_int()
        _check(lib().h1pa_batch_poll(self._h, int(cap), tags, res, C.byref(n)))
        out = []
        for i in range(n.value):
            r = res[i]
            out.append((int(tags[i]), dict(converged=bool(r.converged), iterations=r.iterations,
                                           outer_iterations=r.outer_iterations,
                                           final_norm=r.final_norm,
                                           rel_residual=r.rel_residual, status=r.status)))
            self._keep.pop(int(tags[i]), None)
        return out

    def Info(self):
        info = (C.c_int64 * 6)()
        _check(lib().h1pa_batch_info(self._h, info))
        return dict(live=info[0], running=info[1], bytes=info[2], replays=info[3],
                    captures=info[4], groups=info[5])


class LinearForm:
    """LinearForm with one DomainLFIntegrator, whose coefficient is a
    ConstantCoefficient or a QuadratureFunctionCoefficient holding f at the
    (p+1)^3 Gauss-Legendre points of every element ([ne][q1d^3], x fastest:
    the integrator's default rule, order 2p)."""

    def __init__(self, fes):
        self.fes, self.coeff = fes, None

    def AddDomainIntegrator(self, integ):
        if not isinstance(integ, DomainLFIntegrator):
            raise H1paError(H1PA_EUNSUP, "only DomainLFIntegrator is provided")
        if not isinstance(integ.coeff, (ConstantCoefficient, QuadratureFunctionCoefficient)):
            raise H1paError(H1PA_EUNSUP, "the source is a constant or a quadrature function")
        self.coeff = integ.coeff

    def Assemble(self):
        torch = _torch()
        b = torch.empty(self.fes.ndofs, dtype=torch.float64, device="cuda")
        c = self.coeff
        if isinstance(c, QuadratureFunctionCoefficient):
            # the device reads (p+1)^3 values per element: check the size here
            _ptr(c.data)
            try:
                _check_vec(c.data, (self.fes.p + 1) ** 3 * self.fes.GetNE())
            except ValueError as e:
                raise ValueError(f"QuadratureFunctionCoefficient: {e}") from None
            _check(lib().h1pa_space_domain_lf_qf(self.fes._h, 0.0, _ptr(c.data), _ptr(b),
                                                 _stream()))
        else:
            _check(lib().h1pa_space_domain_lf(self.fes._h, float(c.value), _ptr(b), _stream()))
        return b


class DomainLFIntegrator:
    def __init__(self, coeff):
        self.coeff = coeff
