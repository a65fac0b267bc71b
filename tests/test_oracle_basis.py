"""Pins the oracle's 1D rules, basis and lexicographic dof_map.

Independent sources: the mpmath fixture tests/golden/basis_mp.json (40-digit
roots of P_n / P'_{n-1} and Lagrange interpolation), numpy's leggauss,
polynomial exactness, and MFEM's own hex geometry tables (fem/geom.cpp:
1020-1036) for the dof_map.  Mirrors tests/unit/fem/test_intrules.cpp,
test_poly1d.cpp and test_lexicographic_ordering.cpp of the reference.
"""
import json
import os

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "basis_mp.json")))

HEX_CORNERS = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                        [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
HEX_EDGES = [(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7), (0, 4), (1, 5),
             (2, 6), (3, 7)]
HEX_FACES = [(3, 2, 1, 0), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]


@pytest.mark.parametrize("n", range(1, 12))
def test_gauss_legendre_golden(orc, n):
    x, w = orc.gauss_legendre(n)
    gx = np.array([float(v) for v in GOLD["gl"][str(n)]["x"]])
    gw = np.array([float(v) for v in GOLD["gl"][str(n)]["w"]])
    assert np.max(np.abs(x - gx)) < 2e-16
    assert np.max(np.abs(w - gw)) < 2e-16
    # numpy's Golub-Welsch rule mapped to [0,1]
    nx, nw = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, (nx + 1) / 2, atol=1e-15) and np.allclose(w, nw / 2, atol=1e-15)
    # exact for degree 2n-1
    for k in range(2 * n):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14


@pytest.mark.parametrize("n", range(2, 11))
def test_gauss_lobatto_golden(orc, n):
    x, w = orc.gauss_lobatto(n)
    gx = np.array([float(v) for v in GOLD["gll"][str(n)]["x"]])
    gw = np.array([float(v) for v in GOLD["gll"][str(n)]["w"]])
    assert np.max(np.abs(x - gx)) < 2e-16
    assert np.max(np.abs(w - gw)) < 2e-16
    assert x[0] == 0.0 and x[-1] == 1.0
    for k in range(2 * n - 2):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("dq", [1, 2])
def test_basis_golden(orc, p, dq):
    Q = p + dq
    B, G = orc.basis(p, Q)
    g = GOLD["basis"][f"{p},{Q}"]
    gB = np.array([[float(v) for v in r] for r in g["B"]])
    gG = np.array([[float(v) for v in r] for r in g["G"]])
    assert np.max(np.abs(B - gB)) < 1e-15
    assert np.max(np.abs(G - gG)) <= 1e-14 * max(1.0, np.max(np.abs(gG)))
    # partition of unity / derivative of a constant
    assert np.allclose(B.sum(1), 1.0, atol=1e-14)
    assert np.allclose(G.sum(1), 0.0, atol=1e-12)
    # interpolation of x^k (k <= p) and its derivative is exact
    z, _ = orc.gauss_lobatto(p + 1)
    xq, _ = orc.gauss_legendre(Q)
    for k in range(p + 1):
        assert np.allclose(B @ z ** k, xq ** k, atol=1e-13)
        dk = k * xq ** (k - 1) if k > 0 else 0 * xq
        assert np.allclose(G @ z ** k, dk, atol=1e-11)


def test_legendre_orthogonality(orc):
    """tests/unit/fem/test_poly1d.cpp: int_0^1 P_i P_j = delta_ij / (2i+1)."""
    x, w = orc.gauss_legendre(10)
    P = [np.polynomial.legendre.Legendre.basis(i)(2 * x - 1) for i in range(8)]
    for i in range(8):
        for j in range(8):
            v = np.dot(w, P[i] * P[j])
            assert abs(v - (1.0 / (2 * i + 1) if i == j else 0.0)) < 1e-14


@pytest.mark.parametrize("p", range(1, 9))
def test_dof_map_matches_hex_geometry(orc, p):
    """Each native DOF of H1_HexahedronElement sits on the vertex / edge / face
    the hex tables say, with edge DOFs ordered from the edge's first vertex and
    face DOF (i, j) at v0 + i/p (v1 - v0) + j/p (v3 - v0) of FaceVert."""
    p1 = p + 1
    dm = orc.dof_map(p)
    assert sorted(dm.tolist()) == list(range(p1 ** 3))
    lex_xyz = np.array([[i, j, k] for k in range(p1) for j in range(p1) for i in range(p1)])
    native_xyz = np.zeros_like(lex_xyz)
    native_xyz[dm] = lex_xyz
    C = HEX_CORNERS * p
    for v in range(8):
        assert np.array_equal(native_xyz[v], C[v])
    o = 8
    for a, b in HEX_EDGES:
        for i in range(1, p):
            assert np.array_equal(native_xyz[o], C[a] + i * (C[b] - C[a]) // p)
            o += 1
    for f in HEX_FACES:
        v0, v1, v3 = C[f[0]], C[f[1]], C[f[3]]
        for j in range(1, p):
            for i in range(1, p):
                assert np.array_equal(native_xyz[o], v0 + (i * (v1 - v0) + j * (v3 - v0)) // p)
                o += 1
    for k in range(1, p):
        for j in range(1, p):
            for i in range(1, p):
                assert np.array_equal(native_xyz[o], [i, j, k])
                o += 1
    assert o == p1 ** 3
