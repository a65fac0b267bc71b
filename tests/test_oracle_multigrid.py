"""The oracle's p-refinement transfer (TensorProductPRefinementTransferOperator,
fem/transfer.cpp:2619-2990) and p-multigrid cycle (MultigridBase::Cycle,
fem/multigrid.cpp:106-240, levels as examples/ex26.cpp:42-110), checked by
what the mathematics fixes:

  - prolongation interpolates: a field that is a polynomial of degree <= p_low
    in each coordinate (affine mesh) prolongs to its exact high-order nodal
    values, and a constant stays constant on any mesh;
  - restriction is the transpose of prolongation: (P x, y) = (x, P^T y);
  - one cycle equals the same cycle written with dense matrices (each level's
    constrained operator, the transfers and the Chebyshev polynomial
    assembled column by column, the coarse solve exact);
  - CG preconditioned by the cycle needs far fewer iterations than Jacobi-CG
    and reaches the same solution.
"""
import numpy as np
import pytest


def _dof_coords(orc, P):
    """Physical coordinates of P's L-DOFs (affine mesh): the order-p GLL
    nodes of every element, scattered through the gather map."""
    X = P.mesh.nodes_curved(P.xyz, P.p, 0.0).reshape(P.ne, 3, P.D ** 3)
    c = np.zeros((P.ndofs, 3))
    g = P.gather_map.reshape(P.ne, P.D ** 3)
    for e in range(P.ne):
        c[g[e]] = X[e].T
    return c


@pytest.mark.parametrize("pl,ph", [(1, 2), (2, 4), (3, 5), (1, 8)])
def test_prolongation_interpolates(orc, pl, ph):
    lo = orc.Problem(2, pl)
    hi = orc.Problem(2, ph)
    T = orc.Transfer(lo, hi)
    cl, ch = _dof_coords(orc, lo), _dof_coords(orc, hi)

    def f(c):
        x, y, z = c[:, 0], c[:, 1], c[:, 2]
        return 1.5 - x ** pl * y + 0.25 * z ** pl * x + (y * z) ** min(pl, 2)

    assert np.allclose(T.mult(f(cl)), f(ch), rtol=0, atol=1e-13)
    # a constant on a perturbed mesh (partition of unity of the low basis)
    lo2 = orc.Problem(3, pl, perturb=0.2, seed=4)
    hi2 = orc.Problem(3, ph, perturb=0.2, seed=4)
    y = orc.Transfer(lo2, hi2).mult(np.full(lo2.ndofs, 2.5))
    assert np.allclose(y, 2.5, rtol=0, atol=1e-13)


@pytest.mark.parametrize("pl,ph", [(1, 3), (2, 4), (4, 8)])
def test_restriction_is_the_transpose(orc, pl, ph):
    lo = orc.Problem(3, pl, perturb=0.2, seed=7)
    hi = orc.Problem(3, ph, perturb=0.2, seed=7)
    T = orc.Transfer(lo, hi)
    rng = np.random.default_rng(pl + ph)
    x, y = rng.uniform(-1, 1, lo.ndofs), rng.uniform(-1, 1, hi.ndofs)
    a, b = T.mult(x) @ y, x @ T.mult_transpose(y)
    assert abs(a - b) <= 1e-12 * abs(a)


def _dense(n, op):
    return np.column_stack([op(np.eye(n)[:, j]) for j in range(n)])


def test_cycle_equals_dense_cycle(orc):
    orders = (1, 2, 4)
    levels = [orc.Problem(2, p, perturb=0.15, seed=3, coeff="sincos") for p in orders]
    mg = orc.PMultigrid(levels, smoother_order=2, coarse_rel_tol=1e-15, coarse_max_iter=1000)
    A = [_dense(P.ndofs, P.constrained_mult) for P in levels]
    Pm = [_dense(levels[l].ndofs, mg.P[l].mult) for l in range(2)]
    Rm = [_dense(levels[l + 1].ndofs, mg.P[l].mult_transpose) for l in range(2)]
    for l in range(2):
        assert np.allclose(Rm[l], Pm[l].T, rtol=0, atol=1e-13)
    S = [None]
    for l in (1, 2):
        Dinv = np.diag(1.0 / mg.diag[l])
        for i in levels[l].ess:
            Dinv[i, i] = 1.0
        c = mg.coeffs[l]
        K = Dinv @ A[l]
        term = Dinv.copy()
        Sl = c[0] * term
        for k in range(1, len(c)):
            term = K @ term
            Sl = Sl + c[k] * term
        S.append(Sl)

    def cycle(l, x):
        if l == 0:
            return np.linalg.solve(A[0], x)
        y = S[l] @ x
        yc = cycle(l - 1, Pm[l - 1].T @ (x - A[l] @ y))
        y = y + Pm[l - 1] @ yc
        return y + S[l] @ (x - A[l] @ y)

    x = np.random.default_rng(5).uniform(-1, 1, levels[-1].ndofs)
    want = cycle(2, x)
    got = mg.mult(x)
    assert np.linalg.norm(got - want) <= 1e-9 * np.linalg.norm(want)


def test_multigrid_cg_converges_fast(orc):
    n, p = 4, 4
    levels = [orc.Problem(n, q, perturb=0.2, seed=2, coeff="sincos") for q in (1, 2, 4)]
    mg = orc.PMultigrid(levels)
    fine = levels[-1]
    b = np.random.default_rng(1).uniform(-1, 1, fine.ndofs)
    b[fine.ess] = 0.0
    x, it, conv = mg.pcg(b, rel_tol=1e-10)
    xj, info = fine.pcg(b, rel_tol=1e-10)
    assert conv and info["converged"]
    assert it * 3 < info["iterations"], (it, info["iterations"])
    assert np.linalg.norm(x - xj) <= 1e-7 * np.linalg.norm(xj)
    r = b - fine.constrained_mult(x)
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b)
