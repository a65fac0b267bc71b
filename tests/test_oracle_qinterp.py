"""Oracle: QuadratureInterpolator::Mult (fem/quadinterpolator.cpp:226-341,
fem/qinterp/eval.hpp, grad.hpp:233-375, det.hpp), pinned by closed forms.

- A polynomial field of degree <= p per coordinate on an affine mesh is
  represented exactly: values at the Gauss-Legendre points equal u(x_q),
  physical derivatives grad u(x_q), reference derivatives J^T grad u.
- The mesh nodes as a 3-component field: values are the physical points,
  reference derivatives the Jacobian and determinants det J of the geometric
  factors (GeometricFactors::Compute, mesh/mesh.cpp:15445-15497), on a
  perturbed mesh.
- The byVDIM layout is the byNODES one transposed.
"""
import numpy as np
import pytest

import oracle

F = oracle.QI_VALUES | oracle.QI_DERIVATIVES


def _u(x, y, z):
    return 1.0 + x + 2.0 * y * y - 3.0 * x * z + y * z * z


def _grad_u(x, y, z):
    return np.stack([1.0 - 3.0 * z, 4.0 * y + z * z, -3.0 * x + 2.0 * y * z])


def _evec(P, p, fn):
    """E-vector [D^3][1][ne] of fn sampled at the element DOF nodes (GLL
    points mapped by the trilinear element map)."""
    D = p + 1
    xg, _ = oracle.gauss_lobatto(D)
    ne = P.ne
    X = P.X.reshape(ne, 3, 2, 2, 2)  # [e][c][vz][vy][vx]
    lz = np.stack([1 - xg, xg])  # (2, D) trilinear weights per direction
    pts = np.einsum("ecijk,iz,jy,kx->eczyx", X, lz, lz, lz)  # [e][c][dz][dy][dx]
    return fn(pts[:, 0], pts[:, 1], pts[:, 2]).reshape(ne, 1, D, D, D).ravel()


@pytest.mark.parametrize("p,Q", [(2, 3), (3, 4), (3, 6), (4, 5)])
def test_polynomial_field_exact(p, Q):
    P = oracle.Problem(2, p, Q=Q)  # unperturbed: affine elements
    ne, nq = P.ne, Q ** 3
    xE = _evec(P, p, _u)
    Pq = P.P.reshape(ne, 3, nq)  # physical points [e][c][q]
    xs, ys, zs = Pq[:, 0], Pq[:, 1], Pq[:, 2]
    val, der, _ = oracle.qinterp(p, Q, 1, ne, xE, F)
    assert np.allclose(val.reshape(ne, nq), _u(xs, ys, zs), rtol=0, atol=1e-12)
    g = _grad_u(xs, ys, zs)  # [r][e][q]
    _, dph, _ = oracle.qinterp(p, Q, 1, ne, xE, oracle.QI_PHYSICAL_DERIVATIVES, J=P.J)
    dph = dph.reshape(ne, 3, nq)  # [e][d][q] (vdim 1)
    assert np.allclose(dph, g.transpose(1, 0, 2), rtol=0, atol=1e-11)
    # reference derivatives: h grad u on the unit-cube mesh with h = 1/2
    assert np.allclose(der.reshape(ne, 3, nq), 0.5 * g.transpose(1, 0, 2), rtol=0, atol=1e-11)


@pytest.mark.parametrize("Q", [2, 3, 4])
def test_nodes_field_gives_geometric_factors(Q):
    P = oracle.Problem(3, 1, Q=Q, perturb=0.2, seed=5)
    ne, nq = P.ne, Q ** 3
    flags = oracle.QI_VALUES | oracle.QI_DERIVATIVES | oracle.QI_DETERMINANTS
    val, der, det = oracle.qinterp(1, Q, 3, ne, P.X, flags)
    J, Pp = oracle.geom(ne, 1, Q, P.X)
    assert np.allclose(val, Pp, rtol=0, atol=1e-14)
    assert np.allclose(der, J, rtol=0, atol=1e-13)
    Jm = J.reshape(ne, 3, 3, nq).transpose(0, 3, 2, 1)  # [e][q][r][c]
    assert np.allclose(det, np.linalg.det(Jm).ravel(), rtol=1e-13, atol=0)
    assert det.min() > 0


def test_by_vdim_is_the_transpose():
    p, Q, vdim = 2, 4, 3
    P = oracle.Problem(2, p, Q=Q, perturb=0.1, seed=2)
    ne, nq, D = P.ne, Q ** 3, p + 1
    xE = np.random.default_rng(0).uniform(-1, 1, D ** 3 * vdim * ne)
    v0, d0, _ = oracle.qinterp(p, Q, vdim, ne, xE, F)
    v1, d1, _ = oracle.qinterp(p, Q, vdim, ne, xE, F, by_vdim=True)
    assert np.array_equal(v1, v0.reshape(ne, vdim, nq).transpose(0, 2, 1).ravel())
    assert np.array_equal(d1, d0.reshape(ne, 3, vdim, nq).transpose(0, 3, 1, 2).ravel())


def test_curved_order8_nodes_at_q11():
    """Order-8 curved nodes at 11^3 points (beyond the largest rule of the
    GPU interpolator): interpolated nodes reproduce the geometric factors."""
    P = oracle.Problem(2, 8, curve=0.04)
    ne, nq, Q = P.ne, 11 ** 3, 11
    flags = oracle.QI_VALUES | oracle.QI_DERIVATIVES | oracle.QI_DETERMINANTS
    val, der, det = oracle.qinterp(8, Q, 3, ne, P.X, flags)
    J, Pp = oracle.geom(ne, 8, Q, P.X)
    assert np.allclose(val, Pp, rtol=0, atol=1e-13)
    assert np.allclose(der, J, rtol=0, atol=1e-12)
    Jm = J.reshape(ne, 3, 3, nq).transpose(0, 3, 2, 1)
    assert np.allclose(det, np.linalg.det(Jm).ravel(), rtol=1e-12, atol=0)
    with pytest.raises(ValueError):
        oracle.geom(ne, 8, 12, P.X)
