"""Oracle: the seeded log-normal (Fourier) coefficient of the batched and
served workloads, pinned by what the mathematics fixes.

  - zero amplitude gives k = 1 exactly;
  - one mode equals the closed form exp(a cos(2 pi k.x + phi)) (numpy);
  - modes add in the exponent: the field of two modes is the product of the
    two one-mode fields;
  - the seeded modes are integer wave vectors in [-3,3]^3 without the zero
    vector, phases in [0, 2 pi), amplitude 0.5/sqrt(8); k stays within
    e^(+-8a) = e^(+-sqrt 2) and reaches most of that range;
  - the draw is deterministic per seed and differs between seeds.
"""
import numpy as np
import pytest


def _points(rng, ne, nq):
    return rng.uniform(0, 1, (ne, 3, nq))


def test_zero_amplitude_is_one(orc):
    rng = np.random.default_rng(0)
    P = _points(rng, 5, 27)
    kv, phi, _ = orc.fourier_modes(7)
    k = orc.coeff_fourier(3, 5, P, modes=(kv, phi, 0.0))
    assert np.array_equal(k, np.ones_like(k))


@pytest.mark.parametrize("kvec,phi,a", [((1, 0, 0), 0.0, 0.5), ((-2, 3, 1), 1.3, 0.2),
                                        ((0, 0, -3), 5.9, 1.0)])
def test_single_mode_closed_form(orc, kvec, phi, a):
    rng = np.random.default_rng(1)
    ne, Q = 4, 4
    P = _points(rng, ne, Q ** 3)
    k = orc.coeff_fourier(Q, ne, P, modes=(np.array([kvec]), np.array([phi]), a))
    x, y, z = P[:, 0, :], P[:, 1, :], P[:, 2, :]
    ref = np.exp(a * np.cos(2 * np.pi * (kvec[0] * x + kvec[1] * y + kvec[2] * z) + phi))
    assert np.allclose(k, ref.ravel(), rtol=1e-14, atol=0)


def test_modes_multiply(orc):
    rng = np.random.default_rng(2)
    ne, Q = 3, 5
    P = _points(rng, ne, Q ** 3)
    kv = np.array([[1, -1, 2], [0, 3, -1]])
    phi = np.array([0.4, 2.2])
    both = orc.coeff_fourier(Q, ne, P, modes=(kv, phi, 0.3))
    k0 = orc.coeff_fourier(Q, ne, P, modes=(kv[:1], phi[:1], 0.3))
    k1 = orc.coeff_fourier(Q, ne, P, modes=(kv[1:], phi[1:], 0.3))
    assert np.allclose(both, k0 * k1, rtol=1e-14, atol=0)


@pytest.mark.parametrize("seed", [0, 1, 12345, 2 ** 63 + 5])
def test_seeded_modes_and_bounds(orc, seed):
    kv, phi, a = orc.fourier_modes(seed)
    assert kv.shape == (8, 3) and np.all(np.abs(kv) <= 3)
    assert np.all(np.any(kv != 0, axis=1))
    assert np.all((phi >= 0) & (phi < 2 * np.pi))
    assert a == pytest.approx(0.5 / np.sqrt(8), rel=0, abs=1e-17)
    rng = np.random.default_rng(seed % 1000)
    ne, Q = 64, 6
    P = _points(rng, ne, Q ** 3)
    k = orc.coeff_fourier(Q, ne, P, seed=seed)
    lo, hi = np.exp(-8 * a), np.exp(8 * a)
    assert np.all(k >= lo) and np.all(k <= hi)
    g = np.log(k)
    # RMS 0.5/sqrt(2) = 0.354 for 8 distinct modes (a repeated wave vector
    # adds coherently); a dropped factor or mode would leave this band
    assert 0.25 < g.std() < 0.5
    # the same seed draws the same field; another seed another one
    assert np.array_equal(k, orc.coeff_fourier(Q, ne, P, seed=seed))
    assert not np.array_equal(kv, orc.fourier_modes(seed + 1)[0])


def test_fourier_problem_uses_field(orc):
    """Problem(coeff=('fourier', s)) feeds that field to the PA setup: on an
    undeformed Cartesian mesh D = W k h I (closed form of the setup)."""
    P = orc.Problem(2, 2, coeff=("fourier", 3))
    Q, ne = P.Q, P.ne
    k = orc.coeff_fourier(Q, ne, P.P, seed=3)
    D = P.Dd.reshape(ne, 6, Q ** 3)
    h = 0.5
    W = P.W
    for e in range(ne):
        assert np.allclose(D[e, 0], W * k[e * Q ** 3:(e + 1) * Q ** 3] * h, rtol=1e-14)
        assert np.allclose(D[e, 1], 0.0, atol=1e-17)
