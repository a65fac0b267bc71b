"""Oracle: OperatorChebyshevSmoother, PowerMethod and Vector::Randomize
(linalg/solvers.cpp:455-656, linalg/operator.cpp:886-935,
linalg/vector.cpp:955-966), pinned by closed forms and brute force.

- glibc rand: the restated generator against the C library's own srand/rand
  (run through ctypes) and its widely published first outputs for seed 1.
- Coefficients: 1 - lambda p(lambda) must be the scaled Chebyshev polynomial
  T_m((theta - lambda)/delta) / T_m(theta/delta) (a dropped term or a wrong
  sign breaks the identity).
- Power method: converges to the largest eigenvalue of the dense generalized
  problem A v = lambda D v; 10 steps give a lower bound.
- Chebyshev Mult: the dense polynomial sum_k c_k (D^-1 A)^k D^-1 x.
- Chebyshev-PCG: the dense solve of the constrained system.
"""
import ctypes
import ctypes.util

import numpy as np
import pytest
import scipy.linalg
from numpy.polynomial import chebyshev as Ch

import oracle


def _libc():
    name = ctypes.util.find_library("c")
    if not name:
        pytest.skip("no C library to compare with")
    return ctypes.CDLL(name)


def test_glibc_rand_known_answers():
    assert list(oracle.glibc_rand(1, 5)) == [1804289383, 846930886, 1681692777, 1714636915,
                                             1957747793]
    # seed 0 is taken as 1
    assert list(oracle.glibc_rand(0, 3)) == [1804289383, 846930886, 1681692777]


@pytest.mark.parametrize("seed", [1, 2, 12345, 2**31 + 5, 4000000000])
def test_glibc_rand_matches_libc(seed):
    libc = _libc()
    libc.srand.argtypes = [ctypes.c_uint]
    libc.srand(seed)
    ref = [libc.rand() for _ in range(5000)]
    assert list(oracle.glibc_rand(seed, 5000)) == ref


def test_randomize_range():
    v = oracle.randomize(12345, 10000)
    assert v.min() >= 0.0 and v.max() < 1.0
    assert v[0] == oracle.glibc_rand(12345, 1)[0] / 2.0**31


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_chebyshev_coefficients_are_the_chebyshev_residual(order):
    for lam_max in (0.37, 1.0, 2.9, 11.0):
        c = oracle.chebyshev_coeffs(order, lam_max)
        upper, lower = 1.2 * lam_max, 0.3 * lam_max
        th, de = 0.5 * (upper + lower), 0.5 * (upper - lower)
        lam = np.linspace(0.0, upper, 41)
        q = 1.0 - lam * sum(c[k] * lam ** k for k in range(order))
        t = [0] * order + [1]
        ref = Ch.chebval((th - lam) / de, t) / Ch.chebval(th / de, t)
        assert np.max(np.abs(q - ref)) < 1e-12


def test_chebyshev_order_6_rejected():
    with pytest.raises(ValueError):
        oracle.chebyshev_coeffs(6, 1.0)


def _dense(P):
    n = P.ndofs
    A = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = P.constrained_mult(e)
    d = P.diagonal()
    d[P.ess] = 1.0
    return A, d


@pytest.fixture(scope="module")
def small():
    return oracle.Problem(3, 2, perturb=0.2, seed=3, coeff="sincos")


def test_power_method_converges_to_largest_eigenvalue(small):
    A, d = _dense(small)
    lam = scipy.linalg.eigh(0.5 * (A + A.T), np.diag(d), eigvals_only=True)
    e_long, _ = small.power_method(oracle.randomize(12345, small.ndofs), steps=5000, tol=1e-15)
    assert abs(e_long - lam[-1]) <= 1e-8 * lam[-1]
    e10, _ = small.chebyshev(2)
    assert e10 <= lam[-1] * (1 + 1e-12)
    assert e10 > 0.5 * lam[-1]


@pytest.mark.parametrize("order", [1, 2, 3, 5])
def test_chebyshev_mult_is_the_dense_polynomial(small, order):
    A, d = _dense(small)
    eig, c = small.chebyshev(order)
    x = np.random.default_rng(order).uniform(-1, 1, small.ndofs)
    Dinv = 1.0 / d
    t = Dinv * x
    y = c[0] * t
    for k in range(1, order):
        t = Dinv * (A @ t)
        y = y + c[k] * t
    got = small.chebyshev_mult(c, x)
    assert np.linalg.norm(got - y) <= 1e-12 * np.linalg.norm(y)


@pytest.mark.parametrize("order", [2, 3])
def test_pcg_chebyshev_solves_the_system(small, order):
    A, _ = _dense(small)
    b = np.random.default_rng(9).uniform(-1, 1, small.ndofs)
    _, B = small.form_linear_system(np.zeros(small.ndofs), b)
    eig, c = small.chebyshev(order)
    x, st = small.pcg_chebyshev(c, B, rel_tol=1e-12, max_iter=500)
    xj, sj = small.pcg(B, rel_tol=1e-12, max_iter=500)
    ref = np.linalg.solve(A, B)
    assert st["converged"]
    assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)
    assert st["iterations"] < sj["iterations"]
