"""GPU: QuadratureInterpolator::Mult against the oracle
(fem/quadinterpolator.cpp:226-341).

Bars: values, reference and physical derivatives and determinants within
1e-13 relative (2-norm) of the oracle's on the same E-vector and mesh, for
both Q-vector layouts; an affine-mesh polynomial field of degree <= p is
reproduced to 1e-11.
"""
import numpy as np
import pytest

from test_gpu_parity import Pair
from test_oracle_qinterp import _evec, _grad_u, _u

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _run(S, torch, q1d, vdim, flags, layout, xE):
    M = __import__("mfem_amd")
    qi = S.fes.GetQuadratureInterpolator(q1d, vdim)
    qi.SetOutputLayout(layout)
    ne, nq = S.P.ne, q1d ** 3
    f64 = dict(dtype=torch.float64, device="cuda")
    val = torch.full((nq * vdim * ne,), np.nan, **f64) if flags & M.QI_VALUES else None
    der = (torch.full((nq * vdim * 3 * ne,), np.nan, **f64)
           if flags & (M.QI_DERIVATIVES | M.QI_PHYSICAL_DERIVATIVES) else None)
    det = torch.full((nq * ne,), np.nan, **f64) if flags & M.QI_DETERMINANTS else None
    qi.Mult(S.dev(xE), flags, val, der, det)
    return tuple(None if t is None else S.host(t) for t in (val, der, det))


CASES = [(1, 2), (1, 3), (2, 3), (2, 5), (3, 5), (4, 6), (5, 7), (6, 8), (6, 9), (7, 10),
         (8, 10)]


@pytest.mark.parametrize("p,q1d", CASES)
@pytest.mark.parametrize("layout", [0, 1])
def test_qinterp_matches_oracle(orc, h1pa, torch_cuda, p, q1d, layout):
    M = h1pa
    curve = 0.04 if p >= 2 else 0.0
    S = Pair(orc, h1pa, torch_cuda, (3, 2, 2), p, perturb=0.0 if curve else 0.15, seed=p,
             curve=curve)
    ne, D = S.P.ne, p + 1
    J, _ = orc.geom(ne, S.P.pg, q1d, S.P.X)
    for vdim in (1, 3):
        xE = np.random.default_rng(10 * p + vdim).uniform(-1, 1, D ** 3 * vdim * ne)
        for flags in (M.QI_VALUES | M.QI_DERIVATIVES, M.QI_PHYSICAL_DERIVATIVES,
                      M.QI_VALUES | M.QI_DETERMINANTS):
            if flags & M.QI_DETERMINANTS and vdim != 3:
                continue
            got = _run(S, torch_cuda, q1d, vdim, flags, layout, xE)
            ref = orc.qinterp(p, q1d, vdim, ne, xE, flags, by_vdim=bool(layout), J=J)
            for name, g, r in zip(("val", "der", "det"), got, ref):
                if r is None:
                    assert g is None
                    continue
                err = _rel(g, r)
                print(f"[qi] p={p} q1d={q1d} vdim={vdim} flags={flags} layout={layout} "
                      f"{name} err={err:.2e}", flush=True)
                assert err < 1e-13


@pytest.mark.parametrize("p,q1d", [(2, 3), (3, 6), (4, 5)])
def test_polynomial_field_exact_on_gpu(orc, h1pa, torch_cuda, p, q1d):
    M = h1pa
    S = Pair(orc, h1pa, torch_cuda, (2, 2, 2), p)  # the form keeps its default rule
    P = orc.Problem(2, p, Q=q1d)
    ne, nq = P.ne, q1d ** 3
    xE = _evec(P, p, _u)
    val, dph, _ = _run(S, torch_cuda, q1d, 1, M.QI_VALUES | M.QI_PHYSICAL_DERIVATIVES, 0, xE)
    Pq = P.P.reshape(ne, 3, nq)
    xs, ys, zs = Pq[:, 0], Pq[:, 1], Pq[:, 2]
    assert np.allclose(val.reshape(ne, nq), _u(xs, ys, zs), rtol=0, atol=1e-11)
    assert np.allclose(dph.reshape(ne, 3, nq), _grad_u(xs, ys, zs).transpose(1, 0, 2), rtol=0,
                       atol=1e-10)


def test_qinterp_argument_errors(orc, h1pa, torch_cuda):
    M = h1pa
    S = Pair(orc, h1pa, torch_cuda, (2, 2, 2), 2)
    ne = S.P.ne
    xE = np.zeros(27 * ne)
    with pytest.raises(M.H1paError):  # q1d outside [p+1, min(p+3, 10)]
        _run(S, torch_cuda, 6, 1, M.QI_VALUES, 0, xE)
    with pytest.raises(M.H1paError):
        _run(S, torch_cuda, 2, 1, M.QI_VALUES, 0, xE)
    with pytest.raises(M.H1paError):  # both derivative kinds
        _run(S, torch_cuda, 3, 1, M.QI_DERIVATIVES | M.QI_PHYSICAL_DERIVATIVES, 0, xE)
    with pytest.raises(M.H1paError):  # determinants need vdim = 3
        _run(S, torch_cuda, 3, 1, M.QI_DETERMINANTS, 0, xE)
