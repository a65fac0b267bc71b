"""Python view of the CPU oracle (oracle/h1pa_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg are the only users.  The product (mfem_amd/) never imports
this module.

The oracle restates MFEM's H1 partial-assembly path on the CPU; every C
function cites the MFEM file:line it follows.  `Problem` strings the pieces
together the way MFEM's ex1-style driver does (mesh -> H1 space ->
ElementRestriction -> geometric factors -> PA setup -> apply / diagonal /
constrained Jacobi-PCG).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libh1pa_oracle.so")
_lib = None

_i = C.c_int
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)


class _Mesh(C.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("nv", _i), ("ne", _i),
                ("nedges", _i), ("nfaces", _i), ("ev", _pi), ("ee", _pi),
                ("eeo", _pi), ("ef", _pi), ("efo", _pi), ("f_elem1", _pi),
                ("f_bdr", C.c_void_p)]


class _PaOp(C.Structure):
    _fields_ = [("D1", _i), ("Q1", _i), ("ne", _i), ("ndofs", _i),
                ("gather_map", _pi), ("offsets", _pi), ("indices", _pi),
                ("B", _pd), ("G", _pd), ("Dd", _pd), ("Mv", _pd),
                ("n_ess", _i), ("ess", _pi)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "libh1pa_oracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        L = _lib
        L.orc_mesh_cartesian.argtypes = [_i, _i, _i, _i, C.POINTER(_Mesh)]
        L.orc_mesh_free.argtypes = [C.POINTER(_Mesh)]
        L.orc_mesh_from_elements.argtypes = [_i, _i, _pi, _i, _pi, C.POINTER(_Mesh)]
        L.orc_mesh_from_elements.restype = _i
        L.orc_grid_sfc3d.argtypes = [_i, _i, _i, _pi]
        L.orc_mesh_vertices.argtypes = [_i, _i, _i, _pd]
        L.orc_perturb_vertices.argtypes = [_i, _i, _i, _d, C.c_ulonglong, _pd]
        L.orc_h1_elem_dofs.argtypes = [C.POINTER(_Mesh), _i, _pi]
        L.orc_h1_elem_dofs.restype = _i
        L.orc_h1_dof_map.argtypes = [_i, _pi]
        L.orc_h1_boundary_dofs.argtypes = [C.POINTER(_Mesh), _i, _pi, _pi, _i]
        L.orc_h1_boundary_dofs.restype = _i
        L.orc_restriction_build.argtypes = [_i, _i, _i, _pi, _pi, _pi, _pi, _pi]
        L.orc_restriction_mult.argtypes = [_i, _i, _pi, _pd, _pd]
        L.orc_restriction_mult_transpose.argtypes = [_i, _i, _pi, _pi, _pd, _pd, _i]
        L.orc_gauss_legendre.argtypes = [_i, _pd, _pd]
        L.orc_gauss_lobatto.argtypes = [_i, _pd, _pd]
        L.orc_basis.argtypes = [_i, _i, _pd, _pd]
        L.orc_basis_at.argtypes = [_i, _i, _pd, _pd, _pd]
        L.orc_tensor_weights.argtypes = [_i, _pd]
        L.orc_nodes_trilinear.argtypes = [C.POINTER(_Mesh), _pd, _pd]
        L.orc_nodes_curved.argtypes = [C.POINTER(_Mesh), _pd, _i, _d, _pd]
        L.orc_geom.argtypes = [_i, _i, _i, _pd, _pd, _pd]
        L.orc_coeff_sincos.argtypes = [_i, _i, _pd, _pd]
        L.orc_fourier_modes.argtypes = [C.c_ulonglong, _pi, _pd, _pd]
        L.orc_coeff_fourier.argtypes = [_i, _pi, _pd, _d, _i, _i, _pd, _pd]
        L.orc_diffusion_setup.argtypes = [_i, _i, _pd, _pd, _pd, _i, _pd]
        L.orc_mass_setup.argtypes = [_i, _i, _pd, _pd, _pd, _i, _pd]
        L.orc_diffusion_apply.argtypes = [_i, _i, _i, _pd, _pd, _pd, _pd, _pd]
        L.orc_mass_apply.argtypes = [_i, _i, _i, _pd, _pd, _pd, _pd]
        L.orc_diffusion_diag.argtypes = [_i, _i, _i, _pd, _pd, _pd, _pd]
        L.orc_mass_diag.argtypes = [_i, _i, _i, _pd, _pd, _pd]
        L.orc_pa_mult.argtypes = [C.POINTER(_PaOp), _pd, _pd]
        L.orc_pa_diagonal.argtypes = [C.POINTER(_PaOp), _pd]
        L.orc_constrained_mult.argtypes = [C.POINTER(_PaOp), _pd, _pd]
        L.orc_form_linear_system.argtypes = [C.POINTER(_PaOp), _pd, _pd, _pd, _pd]
        L.orc_pcg.argtypes = [C.POINTER(_PaOp), _pd, _pd, _pd, _d, _d, _i, _pi, _pd, _pd]
        L.orc_pcg.restype = _i
        L.orc_glibc_rand.argtypes = [C.c_uint, _i, _pi]
        L.orc_randomize.argtypes = [C.c_uint, _i, _pd]
        L.orc_chebyshev_coeffs.argtypes = [_i, _d, _pd]
        L.orc_chebyshev_coeffs.restype = _i
        L.orc_power_method.argtypes = [C.POINTER(_PaOp), _pd, _pd, _i, _d]
        L.orc_power_method.restype = _d
        L.orc_chebyshev_mult.argtypes = [C.POINTER(_PaOp), _pd, _i, _pd, _pd, _pd]
        L.orc_pcg_chebyshev.argtypes = [C.POINTER(_PaOp), _pd, _i, _pd, _pd, _pd, _d, _d, _i, _pi,
                                        _pd, _pd]
        L.orc_pcg_chebyshev.restype = _i
        L.orc_domain_lf.argtypes = [_i, _i, _i, _pd, _d, _pd]
        L.orc_domain_lf_qf.argtypes = [_i, _i, _i, _pd, _d, _pd, _pd]
        L.orc_transfer_basis.argtypes = [_i, _i, _pd]
        L.orc_transfer_prolong.argtypes = [_i, _i, _i, _pi, _pi, _pi, _i, _pd, _pd, _pd]
        L.orc_transfer_restrict.argtypes = [_i, _i, _i, _pi, _pi, _pi, _i, _pd, _pd, _pd]
        L.orc_num_threads.restype = _i
        L.orc_set_num_threads.argtypes = [_i]
    return _lib


def _ip(a):
    return a.ctypes.data_as(_pi)


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ia(n):
    return np.zeros(int(n), dtype=np.int32)


def _da(n):
    return np.zeros(int(n), dtype=np.float64)


# ---------------------------------------------------------------------------
# mesh / space / restriction
# ---------------------------------------------------------------------------
def grid_sfc3d(nx, ny, nz):
    c = _ia(3 * nx * ny * nz)
    lib().orc_grid_sfc3d(nx, ny, nz, _ip(c))
    return c.reshape(-1, 3)


def read_mfem_mesh(text):
    """Mesh::ReadMFEMMesh (mesh/mesh_readers.cpp:40-130) restated for
    'MFEM mesh v1.0' text of hexahedra with a quadrilateral boundary and
    straight vertices: returns (ev [ne, 8], bquads [nb, 4], xyz [nv, 3])."""
    lines = text.splitlines()
    if not lines or lines[0].strip() != "MFEM mesh v1.0":
        raise ValueError("not an MFEM mesh v1.0 file")
    toks = []
    for ln in lines[1:]:
        ln = ln.split("#", 1)[0]
        toks.extend(ln.split())
    it = iter(toks)

    def word(expect=None):
        w = next(it)
        if expect is not None and w != expect:
            raise ValueError(f"expected {expect!r}, got {w!r}")
        return w

    word("dimension")
    if int(word()) != 3:
        raise ValueError("only 3D meshes")
    word("elements")
    ne = int(word())
    ev = np.zeros((ne, 8), np.int32)
    for e in range(ne):
        word()  # attribute
        if int(word()) != 5:
            raise ValueError("only hexahedra")
        ev[e] = [int(word()) for _ in range(8)]
    word("boundary")
    nb = int(word())
    bq = np.zeros((nb, 4), np.int32)
    for b in range(nb):
        word()
        if int(word()) != 3:
            raise ValueError("only quadrilateral boundary elements")
        bq[b] = [int(word()) for _ in range(4)]
    word("vertices")
    nv = int(word())
    sd = word()
    if sd != "3":
        raise ValueError("only straight 3D vertices")
    xyz = np.array([float(word()) for _ in range(3 * nv)]).reshape(nv, 3)
    return ev, bq, xyz


class CartesianMesh:
    """Mesh::MakeCartesian3D(nx, ny, nz, HEXAHEDRON, 1, 1, 1, sfc) topology, or
    (elements=(ev, bquads, nv)) a mesh read from a file."""

    def __init__(self, nx, ny, nz, sfc=True, elements=None):
        m = _Mesh()
        if elements is not None:
            ev, bq, nv = elements
            ev = np.ascontiguousarray(ev, dtype=np.int32).ravel()
            bq = np.ascontiguousarray(bq, dtype=np.int32).ravel()
            rc = lib().orc_mesh_from_elements(int(nv), ev.size // 8, _ip(ev), bq.size // 4,
                                              _ip(bq), C.byref(m))
            if rc != 0:
                raise ValueError(f"orc_mesh_from_elements: {rc}")
        elif lib().orc_mesh_cartesian(nx, ny, nz, 1 if sfc else 0, C.byref(m)) != 0:
            raise MemoryError("orc_mesh_cartesian")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nv, self.ne = m.nv, m.ne
        self.nedges, self.nfaces = m.nedges, m.nfaces
        ne = m.ne
        self.ev = np.ctypeslib.as_array(m.ev, (ne * 8,)).copy().reshape(ne, 8)
        self.ee = np.ctypeslib.as_array(m.ee, (ne * 12,)).copy().reshape(ne, 12)
        self.eeo = np.ctypeslib.as_array(m.eeo, (ne * 12,)).copy().reshape(ne, 12)
        self.ef = np.ctypeslib.as_array(m.ef, (ne * 6,)).copy().reshape(ne, 6)
        self.efo = np.ctypeslib.as_array(m.efo, (ne * 6,)).copy().reshape(ne, 6)
        self.f_elem1 = np.ctypeslib.as_array(m.f_elem1, (m.nfaces,)).copy()
        self._m = m

    def __del__(self):
        try:
            lib().orc_mesh_free(C.byref(self._m))
        except Exception:
            pass

    def vertices(self, perturb=0.0, seed=0):
        xyz = _da(3 * self.nv)
        lib().orc_mesh_vertices(self.nx, self.ny, self.nz, _dp(xyz))
        if perturb:
            lib().orc_perturb_vertices(self.nx, self.ny, self.nz, float(perturb),
                                       C.c_ulonglong(seed), _dp(xyz))
        return xyz.reshape(-1, 3)

    def elem_dofs(self, p):
        nd = (p + 1) ** 3
        ed = _ia(self.ne * nd)
        ndofs = lib().orc_h1_elem_dofs(C.byref(self._m), p, _ip(ed))
        return ed.reshape(self.ne, nd), ndofs

    def boundary_dofs(self, p, elem_dof):
        ed = np.ascontiguousarray(elem_dof, dtype=np.int32).ravel()
        n = lib().orc_h1_boundary_dofs(C.byref(self._m), p, _ip(ed), _ip(_ia(1)), 0)
        ess = _ia(n)
        lib().orc_h1_boundary_dofs(C.byref(self._m), p, _ip(ed), _ip(ess), n)
        return ess

    def nodes_trilinear(self, xyz):
        X = _da(8 * 3 * self.ne)
        lib().orc_nodes_trilinear(C.byref(self._m), _dp(np.ascontiguousarray(xyz).ravel()), _dp(X))
        return X

    def nodes_curved(self, xyz, pg, amp):
        X = _da((pg + 1) ** 3 * 3 * self.ne)
        lib().orc_nodes_curved(C.byref(self._m), _dp(np.ascontiguousarray(xyz).ravel()), pg,
                               float(amp), _dp(X))
        return X


def dof_map(p):
    dm = _ia((p + 1) ** 3)
    lib().orc_h1_dof_map(p, _ip(dm))
    return dm


def restriction(elem_dof, dmap, ndofs):
    ed = np.ascontiguousarray(elem_dof, dtype=np.int32).ravel()
    ne, nd = elem_dof.shape
    g, o, ix = _ia(ne * nd), _ia(ndofs + 1), _ia(ne * nd)
    lib().orc_restriction_build(ne, nd, ndofs, _ip(ed), _ip(dmap) if dmap is not None else None,
                                _ip(g), _ip(o), _ip(ix))
    return g, o, ix


def restriction_mult(gather_map, nd, x):
    y = _da(gather_map.size)
    lib().orc_restriction_mult(gather_map.size // nd, nd, _ip(gather_map),
                               _dp(np.ascontiguousarray(x, dtype=np.float64)), _dp(y))
    return y


def restriction_mult_transpose(offsets, indices, nd, xe, mode=0, y=None):
    ndofs = offsets.size - 1
    y = _da(ndofs) if y is None else y.copy()
    lib().orc_restriction_mult_transpose(ndofs, nd, _ip(offsets), _ip(indices),
                                         _dp(np.ascontiguousarray(xe, dtype=np.float64)), _dp(y),
                                         mode)
    return y


def boolean_mask(offsets, indices, nE):
    """ElementRestriction::BooleanMask (fem/restriction.cpp:255-290), vdim 1:
    walking each L-DOF's row of `indices` in order, the first E-entry gets 1
    and every later one 0."""
    y = np.full(nE, -1.0)
    for i in range(offsets.size - 1):
        for j in range(offsets[i], offsets[i + 1]):
            ix = int(indices[j])
            y[ix if ix >= 0 else -1 - ix] = 1.0 if j == offsets[i] else 0.0
    return y


def first_entry_mask(offsets, indices, nE):
    """BooleanMask (fem/restriction.cpp:255-290) vectorized: 1.0 at the first
    entry of every non-empty row of the transpose table."""
    m = np.zeros(nE)
    rows = np.nonzero(offsets[1:] > offsets[:-1])[0]
    first = indices[offsets[rows]]
    m[np.where(first >= 0, first, -1 - first)] = 1.0
    return m


class Transfer:
    """TensorProductPRefinementTransferOperator (fem/transfer.cpp:2619-2990)
    between two Problems of one mesh (lo.p < hi.p)."""

    def __init__(self, lo, hi):
        assert lo.ne == hi.ne and lo.p < hi.p
        self.lo, self.hi = lo, hi
        self.mask_h = first_entry_mask(hi.offsets, hi.indices, hi.ne * hi.D ** 3)

    def mult(self, xl):
        lo, hi = self.lo, self.hi
        y = _da(hi.ndofs)
        lib().orc_transfer_prolong(lo.ne, lo.p, hi.p, _ip(lo.gather_map), _ip(hi.offsets),
                                   _ip(hi.indices), hi.ndofs, _dp(self.mask_h),
                                   _dp(np.ascontiguousarray(xl, dtype=np.float64)), _dp(y))
        return y

    def mult_transpose(self, xh):
        lo, hi = self.lo, self.hi
        y = _da(lo.ndofs)
        lib().orc_transfer_restrict(lo.ne, lo.p, hi.p, _ip(hi.gather_map), _ip(lo.offsets),
                                    _ip(lo.indices), lo.ndofs, _dp(self.mask_h),
                                    _dp(np.ascontiguousarray(xh, dtype=np.float64)), _dp(y))
        return y


def cg_preconditioned(P, M, b, rel_tol, abs_tol=0.0, max_iter=2000):
    """CGSolver::Mult (linalg/solvers.cpp:869-1050), iterative_mode = false,
    with preconditioner z = M(r) on P's DIAG_ONE constrained operator (M =
    None: no preconditioner).  Returns (x, iterations, converged)."""
    b = np.ascontiguousarray(b, dtype=np.float64)
    prec = M if M is not None else (lambda v: v.copy())
    r = b.copy()
    x = np.zeros_like(b)
    z = prec(r)
    d = z.copy()
    nom = float(d @ r)
    if not np.isfinite(nom) or nom < 0.0:
        return x, 0, False
    r0 = max(nom * rel_tol * rel_tol, abs_tol * abs_tol)
    if nom <= r0:
        return x, 0, True
    z = P.constrained_mult(d)
    den = float(z @ d)
    if den <= 0.0 and den == 0.0:
        return x, 0, False
    i = 1
    while True:
        alpha = nom / den
        x += alpha * d
        r -= alpha * z
        z = prec(r)
        betanom = float(r @ z)
        if not np.isfinite(betanom) or betanom < 0.0:
            return x, i, False
        if betanom <= r0:
            return x, i, True
        i += 1
        if i > max_iter:
            return x, max_iter, False
        beta = betanom / nom
        d = z + beta * d
        z = P.constrained_mult(d)
        den = float(d @ z)
        if den <= 0.0 and den == 0.0:
            return x, i, False
        nom = betanom


class PMultigrid:
    """MultigridBase::Mult / Cycle / SmoothingStep (fem/multigrid.cpp:106-240)
    over Problems of one mesh with increasing orders, set up as ex26's
    DiffusionMultigrid (examples/ex26.cpp:42-110): OperatorChebyshevSmoother
    on levels >= 1 (diagonal = AssembleDiagonal, eigenvalue by the power
    method), CGSolver without a preconditioner on the coarsest level, every
    level constrained on its boundary DOFs (the Problems' ess lists)."""

    def __init__(self, levels, smoother_order=2, pre=1, post=1, w_cycle=False,
                 coarse_rel_tol=1e-2, coarse_max_iter=200, power_iterations=10,
                 power_tolerance=1e-8, power_seed=12345, diags=None, coarse_jacobi=False):
        """diags: the smoothers' diagonals per level (None: each level's own);
        coarse_jacobi: OperatorJacobiSmoother on the coarse CG (ex26: none)."""
        self.levels = levels
        self.order, self.pre, self.post, self.w_cycle = smoother_order, pre, post, w_cycle
        self.coarse_rel_tol, self.coarse_max_iter = coarse_rel_tol, coarse_max_iter
        self.coarse_jacobi = coarse_jacobi
        self.diag, self.coeffs, self.max_eig = [None], [None], [0.0]
        for l, P in enumerate(levels[1:], 1):
            dg = P.diagonal() if diags is None else diags[l]
            e, c = P.chebyshev(smoother_order, power_iterations, power_tolerance, power_seed,
                               diag=dg)
            self.diag.append(dg)
            self.coeffs.append(c)
            self.max_eig.append(e)
        self.P = [Transfer(levels[l], levels[l + 1]) for l in range(len(levels) - 1)]
        self.coarse_iterations = 0

    def _smooth(self, l, v):
        return self.levels[l].chebyshev_mult(self.coeffs[l], v, self.diag[l])

    def _cycle(self, l, x, y):
        P = self.levels[l]
        if l == 0:  # SmoothingStep(0, zero = true): the coarse CGSolver
            if self.coarse_jacobi:
                yc, info = P.pcg(x, self.coarse_rel_tol, 0.0, self.coarse_max_iter)
                it = info["iterations"]
            else:
                yc, it, _ = cg_preconditioned(P, None, x, self.coarse_rel_tol, 0.0,
                                              self.coarse_max_iter)
            self.coarse_iterations += it
            return yc
        for i in range(self.pre):
            if not self.w_cycle and i == 0:
                y = self._smooth(l, x)
            else:
                y = y + self._smooth(l, x - P.constrained_mult(y))
        r = x - P.constrained_mult(y)
        xc = self.P[l - 1].mult_transpose(r)
        yc = self._cycle(l - 1, xc, np.zeros_like(xc))
        if self.w_cycle:
            yc = self._cycle(l - 1, xc, yc)
        y = y + self.P[l - 1].mult(yc)
        for _ in range(self.post):
            y = y + self._smooth(l, x - P.constrained_mult(y))
        return y

    def mult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return self._cycle(len(self.levels) - 1, x, np.zeros_like(x))

    def pcg(self, b, rel_tol=1e-12, abs_tol=0.0, max_iter=2000):
        return cg_preconditioned(self.levels[-1], self.mult, b, rel_tol, abs_tol, max_iter)


# ---------------------------------------------------------------------------
# quadrature / basis / geometry / setup / kernels
# ---------------------------------------------------------------------------
def gauss_legendre(n):
    x, w = _da(n), _da(n)
    lib().orc_gauss_legendre(n, _dp(x), _dp(w))
    return x, w


def gauss_lobatto(n):
    x, w = _da(n), _da(n)
    lib().orc_gauss_lobatto(n, _dp(x), _dp(w))
    return x, w


def basis(p, Q):
    """B, G as (Q, D) arrays: B[q, d] = l_d(x_q) (MFEM: B[q + Q*d])."""
    B, G = _da(Q * (p + 1)), _da(Q * (p + 1))
    lib().orc_basis(p, Q, _dp(B), _dp(G))
    return B.reshape(p + 1, Q).T.copy(), G.reshape(p + 1, Q).T.copy()


def basis_at(p, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.size
    B, G = _da(n * (p + 1)), _da(n * (p + 1))
    lib().orc_basis_at(p, n, _dp(x), _dp(B), _dp(G))
    return B.reshape(p + 1, n).T.copy(), G.reshape(p + 1, n).T.copy()


def _flat_BG(B, G):
    # (Q, D) -> MFEM column-major B[q + Q*d]
    return (np.ascontiguousarray(B.T).ravel(), np.ascontiguousarray(G.T).ravel())


def tensor_weights(Q):
    W = _da(Q ** 3)
    lib().orc_tensor_weights(Q, _dp(W))
    return W


def geom(ne, pg, Q, X):
    if Q > 11 or pg > 10:
        raise ValueError("oracle geometry supports Q <= 11 and node order <= 10")
    nq = Q ** 3
    J, P = _da(nq * 9 * ne), _da(nq * 3 * ne)
    lib().orc_geom(ne, pg, Q, _dp(X), _dp(J), _dp(P))
    return J, P


QI_VALUES, QI_DERIVATIVES, QI_DETERMINANTS, QI_PHYSICAL_DERIVATIVES = 1, 2, 4, 8


def qinterp(p, Q, vdim, ne, xE, flags, by_vdim=False, J=None):
    """QuadratureInterpolator::Mult (fem/quadinterpolator.cpp:226-341) with
    the tensor kernels of fem/qinterp/eval.hpp and grad.hpp:233-375 and the
    determinant of fem/qinterp/det.hpp, restated with numpy contractions.

    xE: lexicographic E-vector [D^3][vdim][ne] (dx fastest).  J: geometric
    factors [nq][3][3][ne] (J(q, r, c, e) = dx_r/dxi_c) for physical
    derivatives, which are J^-T times the reference ones.  Returns (val, der,
    det) flat in MFEM's byNODES or byVDIM Q-vector layout (None where not
    requested)."""
    D, nq = p + 1, Q ** 3
    B, G = basis(p, Q)  # (Q, D)
    X = np.asarray(xE, dtype=np.float64).reshape(ne, vdim, D, D, D)  # [e][c][dz][dy][dx]
    val = der = det = None
    want_der = flags & (QI_DERIVATIVES | QI_PHYSICAL_DERIVATIVES)
    if flags & QI_VALUES:
        v = np.einsum("ecijk,zi,yj,xk->eczyx", X, B, B, B).reshape(ne, vdim, nq)
        val = (v.transpose(0, 2, 1) if by_vdim else v).ravel()
    if want_der or flags & QI_DETERMINANTS:
        g = np.stack([np.einsum("ecijk,zi,yj,xk->eczyx", X, B, B, G),
                      np.einsum("ecijk,zi,yj,xk->eczyx", X, B, G, B),
                      np.einsum("ecijk,zi,yj,xk->eczyx", X, G, B, B)],
                     axis=1).reshape(ne, 3, vdim, nq)  # [e][d][c][q]
        if flags & QI_DETERMINANTS:
            M = g.transpose(0, 3, 2, 1)  # [e][q][c][d]
            det = np.linalg.det(M).ravel()  # [e][q] -> q + nq e
        if want_der:
            if flags & QI_PHYSICAL_DERIVATIVES:
                Jm = np.asarray(J).reshape(ne, 3, 3, nq)  # [e][col][r][q]
                Mt = Jm.transpose(0, 3, 1, 2)  # [e][q][col][r] = J^T
                gq = g.transpose(0, 3, 1, 2)  # [e][q][d][c]
                gp = np.linalg.solve(Mt, gq)  # J^T grad_x = grad_xi
                g = gp.transpose(0, 2, 3, 1)  # [e][d][c][q]
            der = (g.transpose(0, 3, 1, 2) if by_vdim else g).ravel()
    return val, der, det


def fourier_modes(seed):
    """(kv[8][3] int32, phi[8], amplitude) of the seeded log-normal field."""
    kv = np.zeros(24, np.int32)
    phi = _da(8)
    a = C.c_double()
    lib().orc_fourier_modes(int(seed) & ((1 << 64) - 1), _ip(kv), _dp(phi), C.byref(a))
    return kv.reshape(8, 3), phi, a.value


def coeff_fourier(Q, ne, P, seed=None, modes=None):
    """k = exp(sum_m a cos(2 pi k_m . x + phi_m)) at the points P[nq][3][ne]:
    the seeded field (seed) or explicit modes = (kv[m][3], phi[m], a)."""
    kv, phi, a = fourier_modes(seed) if modes is None else modes
    kv = np.ascontiguousarray(kv, dtype=np.int32).reshape(-1, 3)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    C_ = _da(Q ** 3 * ne)
    lib().orc_coeff_fourier(kv.shape[0], _ip(kv), _dp(phi), float(a), Q ** 3, ne,
                            _dp(np.ascontiguousarray(P, dtype=np.float64)), _dp(C_))
    return C_


def coeff_sincos(Q, ne, P):
    C_ = _da(Q ** 3 * ne)
    lib().orc_coeff_sincos(Q ** 3, ne, _dp(P), _dp(C_))
    return C_


def diffusion_setup(Q, ne, W, J, coef):
    coef = np.atleast_1d(np.asarray(coef, dtype=np.float64))
    Dd = _da(Q ** 3 * 6 * ne)
    lib().orc_diffusion_setup(Q, ne, _dp(W), _dp(J), _dp(coef), coef.size, _dp(Dd))
    return Dd


def mass_setup(Q, ne, W, J, coef):
    coef = np.atleast_1d(np.asarray(coef, dtype=np.float64))
    Mv = _da(Q ** 3 * ne)
    lib().orc_mass_setup(Q, ne, _dp(W), _dp(J), _dp(coef), coef.size, _dp(Mv))
    return Mv


def diffusion_apply(p, Q, ne, B, G, Dd, xe, ye=None):
    Bf, Gf = _flat_BG(B, G)
    ye = _da(xe.size) if ye is None else ye.copy()
    lib().orc_diffusion_apply(p + 1, Q, ne, _dp(Bf), _dp(Gf), _dp(Dd),
                              _dp(np.ascontiguousarray(xe)), _dp(ye))
    return ye


def mass_apply(p, Q, ne, B, Mv, xe, ye=None):
    Bf, _ = _flat_BG(B, B)
    ye = _da(xe.size) if ye is None else ye.copy()
    lib().orc_mass_apply(p + 1, Q, ne, _dp(Bf), _dp(Mv), _dp(np.ascontiguousarray(xe)), _dp(ye))
    return ye


def diffusion_diag(p, Q, ne, B, G, Dd):
    Bf, Gf = _flat_BG(B, G)
    ye = _da((p + 1) ** 3 * ne)
    lib().orc_diffusion_diag(p + 1, Q, ne, _dp(Bf), _dp(Gf), _dp(Dd), _dp(ye))
    return ye


def mass_diag(p, Q, ne, B, Mv):
    Bf, _ = _flat_BG(B, B)
    ye = _da((p + 1) ** 3 * ne)
    lib().orc_mass_diag(p + 1, Q, ne, _dp(Bf), _dp(Mv), _dp(ye))
    return ye


def domain_lf(p, ne, pg, X, f=1.0, fq=None):
    """DomainLFIntegrator element vectors with the constant f, or with f
    given at the (p+1)^3 Gauss-Legendre points of every element (fq[q + nq e])."""
    bE = _da((p + 1) ** 3 * ne)
    if fq is None:
        lib().orc_domain_lf(p, ne, pg, _dp(X), float(f), _dp(bE))
    else:
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        assert fq.size == (p + 1) ** 3 * ne
        lib().orc_domain_lf_qf(p, ne, pg, _dp(X), 0.0, _dp(fq), _dp(bE))
    return bE


def glibc_rand(seed, n):
    """srand(seed); n x rand() of glibc's default generator (restated)."""
    o = _ia(n)
    lib().orc_glibc_rand(seed, n, _ip(o))
    return o


def randomize(seed, n):
    """Vector::Randomize(seed): rand() / (RAND_MAX + 1)."""
    v = _da(n)
    lib().orc_randomize(seed, n, _dp(v))
    return v


def chebyshev_coeffs(order, max_eig):
    """OperatorChebyshevSmoother::Setup coefficients (orders 1..5)."""
    c = _da(5)
    if not lib().orc_chebyshev_coeffs(order, max_eig, _dp(c)):
        raise ValueError(f"Chebyshev order {order} not supported")
    return c[:order].copy()


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def round_bf16(a):
    """FP64 -> FP32 (round to nearest even) -> BF16 (round to nearest even on
    the dropped 16 bits), returned as FP64 values.  The storage rule of the
    reduced-precision PA data (no MFEM counterpart)."""
    f = np.ascontiguousarray(a, dtype=np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    nan = np.isnan(f)
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


# ---------------------------------------------------------------------------
# a whole problem, MFEM ex1-style
# ---------------------------------------------------------------------------
class Problem:
    """H1 order-p PA problem on an n^3 (or nx*ny*nz) Cartesian hex mesh.

    perturb: interior vertex noise amplitude as a fraction of h (cfg2: 0.2).
    curve:   amplitude of the cfg3 curving map (geometry of order p) or 0.
    coeff:   'one' | 'sincos' | ('fourier', seed) | ndarray of per-element
             constants.
    integrators: subset of {'diffusion', 'mass'}.
    """

    def __init__(self, n, p, Q=None, perturb=0.0, seed=0, curve=0.0, coeff="one",
                 integrators=("diffusion",), sfc=True, dims=None, mesh_text=None,
                 mesh_arrays=None, nodes=None):
        self.p, self.D = p, p + 1
        self.Q = Q if Q is not None else p + 2
        if mesh_text is not None or mesh_arrays is not None:
            # a mesh file's text, or (ev [ne, 8], bquads [nb, 4], xyz [nv, 3])
            ev, bq, xyz = read_mfem_mesh(mesh_text) if mesh_text is not None else mesh_arrays
            xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
            self.mesh = CartesianMesh(0, 0, 0, elements=(ev, bq, xyz.shape[0]))
            m = self.mesh
            self.ne = m.ne
            self.xyz = xyz
        else:
            nx, ny, nz = dims if dims is not None else (n, n, n)
            self.mesh = CartesianMesh(nx, ny, nz, sfc=sfc)
            m = self.mesh
            self.ne = m.ne
            self.xyz = m.vertices(perturb, seed)
        self.elem_dof, self.ndofs = m.elem_dofs(p)
        self.dof_map = dof_map(p)
        self.gather_map, self.offsets, self.indices = restriction(self.elem_dof, self.dof_map,
                                                                  self.ndofs)
        self.ess = m.boundary_dofs(p, self.elem_dof)
        self.B, self.G = basis(p, self.Q)
        self.W = tensor_weights(self.Q)
        if nodes is not None:  # (order, E-vector [(order+1)^3][3][ne]) given
            self.pg = int(nodes[0])
            self.X = np.ascontiguousarray(nodes[1], dtype=np.float64).ravel()
            assert self.X.size == (self.pg + 1) ** 3 * 3 * self.ne
        elif curve:
            self.pg = p
            self.X = m.nodes_curved(self.xyz, p, curve)
        else:
            self.pg = 1
            self.X = m.nodes_trilinear(self.xyz)
        self.J, self.P = geom(self.ne, self.pg, self.Q, self.X)
        nq = self.Q ** 3
        if isinstance(coeff, tuple) and coeff[0] == "fourier":
            self.C = coeff_fourier(self.Q, self.ne, self.P, seed=coeff[1])
        elif isinstance(coeff, str):
            if coeff == "one":
                self.C = np.ones(1)
            elif coeff == "sincos":
                self.C = coeff_sincos(self.Q, self.ne, self.P)
            else:
                raise ValueError(coeff)
        else:
            pe = np.asarray(coeff, dtype=np.float64)
            assert pe.size == self.ne
            self.C = np.repeat(pe, nq)
        self.Dd = diffusion_setup(self.Q, self.ne, self.W, self.J, self.C) \
            if "diffusion" in integrators else None
        self.Mv = mass_setup(self.Q, self.ne, self.W, self.J, self.C) \
            if "mass" in integrators else None
        Bf, Gf = _flat_BG(self.B, self.G)
        self._Bf, self._Gf = Bf, Gf
        op = _PaOp()
        op.D1, op.Q1, op.ne, op.ndofs = self.D, self.Q, self.ne, self.ndofs
        op.gather_map, op.offsets, op.indices = (_ip(self.gather_map), _ip(self.offsets),
                                                 _ip(self.indices))
        op.B, op.G = _dp(Bf), _dp(Gf)
        op.Dd = _dp(self.Dd) if self.Dd is not None else None
        op.Mv = _dp(self.Mv) if self.Mv is not None else None
        op.n_ess, op.ess = int(self.ess.size), _ip(self.ess)
        self._op = op

    def set_ess(self, ess):
        """Replace the essential (DIAG_ONE) list of the constrained operator
        (ConstrainedOperator's list, linalg/operator.cpp:601-662)."""
        self.ess = np.ascontiguousarray(ess, dtype=np.int32)
        self._op.n_ess, self._op.ess = int(self.ess.size), _ip(self.ess)

    def replace_pa(self, Dd=None, Mv=None):
        """Swap in other PA data of the same shape (e.g. rounded copies)."""
        if Dd is not None:
            assert self.Dd is not None and Dd.size == self.Dd.size
            self.Dd = np.ascontiguousarray(Dd, dtype=np.float64)
            self._op.Dd = _dp(self.Dd)
        if Mv is not None:
            assert self.Mv is not None and Mv.size == self.Mv.size
            self.Mv = np.ascontiguousarray(Mv, dtype=np.float64)
            self._op.Mv = _dp(self.Mv)

    def mult(self, x):
        y = _da(self.ndofs)
        lib().orc_pa_mult(C.byref(self._op), _dp(np.ascontiguousarray(x, dtype=np.float64)),
                          _dp(y))
        return y

    def constrained_mult(self, x):
        y = _da(self.ndofs)
        lib().orc_constrained_mult(C.byref(self._op),
                                   _dp(np.ascontiguousarray(x, dtype=np.float64)), _dp(y))
        return y

    def diagonal(self):
        d = _da(self.ndofs)
        lib().orc_pa_diagonal(C.byref(self._op), _dp(d))
        return d

    def rhs(self, f=1.0, fq=None):
        bE = domain_lf(self.p, self.ne, self.pg, self.X, f, fq)
        return restriction_mult_transpose(self.offsets, self.indices, self.D ** 3, bE, 0)

    def form_linear_system(self, x, b):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        X, Bv = _da(self.ndofs), _da(self.ndofs)
        lib().orc_form_linear_system(C.byref(self._op), _dp(x),
                                     _dp(np.ascontiguousarray(b, dtype=np.float64)), _dp(X),
                                     _dp(Bv))
        return X, Bv

    def power_method(self, v0, steps=10, tol=1e-8, diag=None):
        """PowerMethod::EstimateLargestEigenvalue on diag^-1 A (constrained);
        returns (estimate, final vector)."""
        diag = self.diagonal() if diag is None else diag
        v = np.ascontiguousarray(v0, dtype=np.float64).copy()
        e = lib().orc_power_method(C.byref(self._op), _dp(diag), _dp(v), steps, tol)
        return e, v

    def chebyshev(self, order, power_iterations=10, power_tolerance=1e-8, power_seed=12345,
                  max_eig=None, diag=None):
        """OperatorChebyshevSmoother set-up: (max_eig, coeffs); diag: the
        smoother's diagonal (None: this operator's)."""
        if max_eig is None:
            max_eig, _ = self.power_method(randomize(power_seed, self.ndofs), power_iterations,
                                           power_tolerance, diag=diag)
        return max_eig, chebyshev_coeffs(order, max_eig)

    def chebyshev_mult(self, coeffs, x, diag=None):
        diag = self.diagonal() if diag is None else diag
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        y = _da(self.ndofs)
        lib().orc_chebyshev_mult(C.byref(self._op), _dp(diag), c.size, _dp(c),
                                 _dp(np.ascontiguousarray(x, dtype=np.float64)), _dp(y))
        return y

    def pcg_chebyshev(self, coeffs, b, rel_tol=1e-12, abs_tol=0.0, max_iter=2000, diag=None):
        diag = self.diagonal() if diag is None else diag
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        x = _da(self.ndofs)
        it, fn, n0 = _i(0), _d(0), _d(0)
        conv = lib().orc_pcg_chebyshev(C.byref(self._op), _dp(diag), c.size, _dp(c),
                                       _dp(np.ascontiguousarray(b, dtype=np.float64)), _dp(x),
                                       rel_tol, abs_tol, max_iter, C.byref(it), C.byref(fn),
                                       C.byref(n0))
        return x, dict(converged=bool(conv), iterations=it.value, final_norm=fn.value,
                       initial_norm=n0.value)

    def pcg(self, b, rel_tol=1e-12, abs_tol=0.0, max_iter=2000, diag=None):
        diag = self.diagonal() if diag is None else diag
        x = _da(self.ndofs)
        it, fn, n0 = _i(0), _d(0), _d(0)
        conv = lib().orc_pcg(C.byref(self._op), _dp(diag),
                             _dp(np.ascontiguousarray(b, dtype=np.float64)), _dp(x),
                             rel_tol, abs_tol, max_iter, C.byref(it), C.byref(fn), C.byref(n0))
        return x, dict(converged=bool(conv), iterations=it.value, final_norm=fn.value,
                       initial_norm=n0.value)
