"""Pins the oracle's geometry, PA setup, apply, diagonal, RHS and PCG.

Independent checks (no sum factorization, no adj(J) formula):
  * J from central differences of the explicit trilinear map (exact for a
    trilinear map) and numpy.linalg.inv / det;
  * dense element matrices  K_e = sum_q w k detJ (J^-T grad phi_i).(J^-T grad phi_j),
    M_e = sum_q w k detJ phi_i phi_j  with a full 3D tensor basis;
  * invariants: A 1 = 0, 1'M1 = |Omega| = 1, symmetry, and the exact energy
    |c|^2 of a linear function on a trilinear mesh;
  * PCG: the textbook Jacobi-PCG recurrence with MFEM's stopping rule on the
    dense constrained matrix, and numpy's direct solve.
This mirrors the reference's PA-vs-LEGACY differential tests
(tests/unit/fem/test_pa_kernels.cpp:706-759, test_assembly_levels.cpp,
test_pa_diagonal.cpp, test_operatorjacobismoother.cpp).
"""
import numpy as np
import pytest


def trilinear_map(V, s):
    """V[8][3] (lex corners), s[..., 3] reference points -> physical points."""
    out = np.zeros(s.shape[:-1] + (3,))
    for l in range(8):
        N = np.ones(s.shape[:-1])
        for c in range(3):
            N = N * (s[..., c] if (l >> c) & 1 else 1 - s[..., c])
        out += N[..., None] * V[l]
    return out


def lex_corners(orc, m, xyz, e):
    dm1 = orc.dof_map(1)
    return np.array([xyz[m.ev[e, dm1[l]]] for l in range(8)])


def qpoints(orc, Q):
    x, w = orc.gauss_legendre(Q)
    S = np.array([[x[i], x[j], x[k]] for k in range(Q) for j in range(Q) for i in range(Q)])
    W = np.array([w[i] * w[j] * w[k] for k in range(Q) for j in range(Q) for i in range(Q)])
    return S, W


def fd_jacobian(V, S, h=1e-3):
    J = np.zeros((S.shape[0], 3, 3))
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        J[:, :, c] = (trilinear_map(V, S + e) - trilinear_map(V, S - e)) / (2 * h)
    return J


def tensor_basis(orc, p, S):
    """phi[q, lex], dphi[q, lex, 3] of the GLL Lagrange tensor basis."""
    D = p + 1
    Bx, Gx = orc.basis_at(p, S[:, 0])
    By, Gy = orc.basis_at(p, S[:, 1])
    Bz, Gz = orc.basis_at(p, S[:, 2])
    nq = S.shape[0]
    phi = np.zeros((nq, D ** 3))
    dphi = np.zeros((nq, D ** 3, 3))
    for k in range(D):
        for j in range(D):
            for i in range(D):
                l = i + D * (j + D * k)
                phi[:, l] = Bx[:, i] * By[:, j] * Bz[:, k]
                dphi[:, l, 0] = Gx[:, i] * By[:, j] * Bz[:, k]
                dphi[:, l, 1] = Bx[:, i] * Gy[:, j] * Bz[:, k]
                dphi[:, l, 2] = Bx[:, i] * By[:, j] * Gz[:, k]
    return phi, dphi


def dense_operator(orc, P, kfun=None, mass=False):
    """Assembled global matrix of the problem P (trilinear geometry only)."""
    p, Q = P.p, P.Q
    S, W = qpoints(orc, Q)
    phi, dphi = tensor_basis(orc, p, S)
    n = P.ndofs
    A = np.zeros((n, n))
    D3 = (p + 1) ** 3
    g = P.gather_map.reshape(P.ne, D3)
    for e in range(P.ne):
        V = lex_corners(orc, P.mesh, P.xyz, e)
        J = fd_jacobian(V, S)
        det = np.linalg.det(J)
        k = np.ones(len(S)) if kfun is None else kfun(trilinear_map(V, S))
        if mass:
            Ke = np.einsum("q,qi,qj->ij", W * k * det, phi, phi)
        else:
            Jinv = np.linalg.inv(J)  # [q, c, r]
            grad = np.einsum("qcr,qlc->qlr", Jinv, dphi)  # J^-T grad_ref
            Ke = np.einsum("q,qir,qjr->ij", W * k * det, grad, grad)
        A[np.ix_(g[e], g[e])] += Ke
    return A


def sincos(X):
    return 1 + 0.5 * np.sin(2 * np.pi * X[..., 0]) * np.cos(2 * np.pi * X[..., 1]) * \
        np.sin(2 * np.pi * X[..., 2])


@pytest.mark.parametrize("perturb", [0.0, 0.2])
def test_geometry(orc, perturb):
    P = orc.Problem(3, 2, perturb=perturb, seed=5)
    Q = P.Q
    S, _ = qpoints(orc, Q)
    nq = Q ** 3
    J = P.J.reshape(P.ne, 3, 3, nq)  # [e][c][r][q]
    Pq = P.P.reshape(P.ne, 3, nq)
    for e in range(P.ne):
        V = lex_corners(orc, P.mesh, P.xyz, e)
        Jref = fd_jacobian(V, S)  # [q][r][c]
        assert np.max(np.abs(J[e].transpose(2, 1, 0) - Jref)) < 1e-10
        assert np.max(np.abs(Pq[e].T - trilinear_map(V, S))) < 1e-15
        if perturb == 0.0:
            assert np.allclose(Jref, np.eye(3) / 3, atol=1e-12)


def test_setup_closed_form(orc):
    """Cartesian mesh, k = 1: D = W h I, mass v = W h^3."""
    n = 4
    P = orc.Problem(n, 3, integrators=("diffusion", "mass"))
    nq = P.Q ** 3
    Dd = P.Dd.reshape(P.ne, 6, nq)
    h = 1.0 / n
    for c, want in enumerate([1, 0, 0, 1, 0, 1]):
        assert np.allclose(Dd[:, c, :], want * P.W[None, :] * h, atol=1e-15)
    assert np.allclose(P.Mv.reshape(P.ne, nq), P.W[None, :] * h ** 3, atol=1e-16)


@pytest.mark.parametrize("p,n", [(1, 2), (2, 2), (3, 2), (4, 1), (5, 1), (6, 1)])
@pytest.mark.parametrize("mass", [False, True])
def test_apply_and_diag_vs_dense(orc, p, n, mass):
    P = orc.Problem(n, p, perturb=0.2, seed=3, coeff="sincos",
                    integrators=("mass",) if mass else ("diffusion",))
    A = dense_operator(orc, P, kfun=sincos, mass=mass)
    x = np.random.default_rng(p).uniform(-1, 1, P.ndofs)
    y = P.mult(x)
    ref = A @ x
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    d = P.diagonal()
    assert np.linalg.norm(d - np.diag(A)) <= 1e-13 * np.linalg.norm(np.diag(A))


@pytest.mark.parametrize("p", [1, 2, 3, 4, 6, 8])
def test_invariants(orc, p):
    n = 3 if p <= 4 else 2
    P = orc.Problem(n, p, perturb=0.2, seed=9)
    one = np.ones(P.ndofs)
    assert np.max(np.abs(P.mult(one))) < 1e-12  # A 1 = 0 (no BCs)
    Pm = orc.Problem(n, p, perturb=0.2, seed=9, integrators=("mass",))
    assert abs(one @ Pm.mult(one) - 1.0) < 1e-13  # |Omega| = 1
    rng = np.random.default_rng(1)
    x, y = rng.uniform(-1, 1, P.ndofs), rng.uniform(-1, 1, P.ndofs)
    assert abs(x @ P.mult(y) - y @ P.mult(x)) < 1e-12 * P.ndofs
    # linear u = c.x is in the space; a(u,u) = |c|^2 |Omega| exactly
    X = P.mesh.nodes_curved(P.xyz, p, 0.0).reshape(P.ne, 3, -1)
    u = np.zeros(P.ndofs)
    c = np.array([1.0, 2.0, 3.0])
    u[P.gather_map] = np.einsum("c,ecl->el", c, X).ravel()
    assert abs(u @ P.mult(u) - 14.0) < 1e-11


def test_curved_geometry_converges(orc):
    """Order-p nodes of the curving map: detJ integrates to the volume of the
    mapped cube (= 1, the map fixes the boundary faces only approximately, so
    compare successive orders) and J approaches the analytic Jacobian."""
    amp = 0.05
    tp = 2 * np.pi
    errs = []
    for pg in (2, 4, 6, 8):
        P = orc.Problem(2, pg, curve=amp, integrators=("mass",))
        nq = P.Q ** 3
        J = P.J.reshape(P.ne, 3, 3, nq)
        # analytic map composed with the (affine) grid map: F(x) = x + amp*(...)
        err = 0.0
        for e in range(P.ne):
            V = lex_corners(orc, P.mesh, P.xyz, e)
            S, _ = qpoints(orc, P.Q)
            Xp = trilinear_map(V, S)
            x, y, z = Xp[:, 0], Xp[:, 1], Xp[:, 2]
            dF = np.zeros((len(S), 3, 3))
            dF[:, 0, 0] = 1; dF[:, 1, 1] = 1; dF[:, 2, 2] = 1
            dF[:, 0, 1] = amp * tp * np.cos(tp * y) * np.sin(tp * z)
            dF[:, 0, 2] = amp * tp * np.sin(tp * y) * np.cos(tp * z)
            dF[:, 1, 2] = amp * tp * np.cos(tp * z) * np.sin(tp * x)
            dF[:, 1, 0] = amp * tp * np.sin(tp * z) * np.cos(tp * x)
            dF[:, 2, 0] = amp * tp * np.cos(tp * x) * np.sin(tp * y)
            dF[:, 2, 1] = amp * tp * np.sin(tp * x) * np.cos(tp * y)
            Jref = dF @ fd_jacobian(V, S)
            err = max(err, np.max(np.abs(J[e].transpose(2, 1, 0) - Jref)))
        errs.append(err)
    assert errs[-1] < 1e-4 and errs[0] > errs[1] > errs[2] > errs[3]


def test_rhs_vs_dense(orc):
    P = orc.Problem(2, 3, perturb=0.2, seed=2)
    b = P.rhs(1.0)
    S, W = qpoints(orc, P.p + 1)
    phi, _ = tensor_basis(orc, P.p, S)
    ref = np.zeros(P.ndofs)
    g = P.gather_map.reshape(P.ne, -1)
    for e in range(P.ne):
        V = lex_corners(orc, P.mesh, P.xyz, e)
        det = np.linalg.det(fd_jacobian(V, S))
        np.add.at(ref, g[e], np.einsum("q,ql->l", W * det, phi))
    assert np.linalg.norm(b - ref) < 1e-13 * np.linalg.norm(ref)
    assert abs(b.sum() - 1.0) < 1e-13  # sum of basis = 1 -> integral of f = 1


def _qpoint_coords(orc, P):
    """Physical coordinates [ne][nq][3] of the (p+1)^3 Gauss-Legendre points
    (x fastest) of every element of a trilinear mesh."""
    S, _ = qpoints(orc, P.p + 1)
    return np.array([trilinear_map(lex_corners(orc, P.mesh, P.xyz, e), S) for e in range(P.ne)])


@pytest.mark.parametrize("p", [1, 2, 4])
def test_rhs_per_point_coefficient(orc, p):
    """DomainLFIntegrator with f given at the quadrature points: a constant
    array is the constant rhs, and on the unit cube b.u for u in the space
    is the exact integral of f u (Gauss-Legendre with p+1 points is exact for
    degree 2p+1 per coordinate)."""
    P = orc.Problem(3, p)
    nq = (p + 1) ** 3
    assert np.array_equal(P.rhs(fq=np.full(nq * P.ne, 2.0)), P.rhs(2.0))
    Xq = _qpoint_coords(orc, P)  # [ne][nq][3]
    x, y, z = (Xq[..., c].ravel() for c in range(3))
    one = np.ones(P.ndofs)
    # linear u = x in the space
    Xn = P.mesh.nodes_curved(P.xyz, p, 0.0).reshape(P.ne, 3, -1)
    ux = np.zeros(P.ndofs)
    ux[P.gather_map] = Xn[:, 0, :].ravel()
    cases = [(y, one, 0.5), (x * y * z, one, 0.125), (y, ux, 0.25), (z * z, ux, 1.0 / 6.0)]
    if p >= 2:
        cases.append((x * y ** 2, ux, 1.0 / 9.0))
    for fq, u, want in cases:
        b = P.rhs(fq=fq)
        assert abs(b @ u - want) < 1e-13, (p, want, b @ u)
    # linearity in f
    b1, b2 = P.rhs(fq=y), P.rhs(fq=z * z)
    assert np.allclose(P.rhs(fq=2 * y - 3 * z * z), 2 * b1 - 3 * b2, rtol=0, atol=1e-14)


def textbook_pcg(A, dinv, b, rel_tol, max_iter):
    """Jacobi-PCG with MFEM's (Br, r) stopping rule; returns (x, iterations)."""
    x = np.zeros_like(b)
    r = b.copy()
    z = dinv * r
    d = z.copy()
    nom = nom0 = z @ r
    if nom <= rel_tol ** 2 * nom0:
        return x, 0
    for i in range(1, max_iter + 1):
        Ad = A @ d
        alpha = nom / (d @ Ad)
        x += alpha * d
        r -= alpha * Ad
        z = dinv * r
        betanom = r @ z
        if betanom <= rel_tol ** 2 * nom0:
            return x, i
        d = z + (betanom / nom) * d
        nom = betanom
    return x, max_iter


@pytest.mark.parametrize("p,n", [(1, 4), (2, 3), (3, 2)])
def test_pcg_vs_textbook(orc, p, n):
    P = orc.Problem(n, p, perturb=0.15, seed=4)
    A = dense_operator(orc, P)
    ess = P.ess
    # constrained operator (DIAG_ONE) and RHS as FormLinearSystem builds them
    Ac = A.copy()
    Ac[ess, :] = 0
    Ac[:, ess] = 0
    Ac[ess, ess] = 1
    b = P.rhs(1.0)
    X, Bv = P.form_linear_system(np.zeros(P.ndofs), b)
    Bref = b.copy()
    Bref[ess] = 0
    assert np.allclose(Bv, Bref, atol=1e-15) and np.all(X == 0)
    xr = np.random.default_rng(0).uniform(-1, 1, P.ndofs)
    assert np.linalg.norm(P.constrained_mult(xr) - Ac @ xr) < 1e-13 * np.linalg.norm(Ac @ xr)
    x, st = P.pcg(Bv, rel_tol=1e-10, max_iter=500)
    dinv = 1.0 / np.diag(Ac)
    xt, it = textbook_pcg(Ac, dinv, Bv, 1e-10, 500)
    assert st["converged"] and abs(st["iterations"] - it) <= 1
    xs = np.linalg.solve(Ac, Bv)
    assert np.linalg.norm(x - xs) < 1e-8 * np.linalg.norm(xs)
