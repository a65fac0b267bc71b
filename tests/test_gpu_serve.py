"""GPU: the native serving loop (mfem_amd/h1pa_serve) on a small request mix,
both continuous batching and the serial FP64 baseline, and the device RNG
fill it uses for right-hand sides."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "mfem_amd", "h1pa_serve")

MASK = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & MASK
    return x ^ (x >> 31)


def test_vector_randomize_bit_exact(h1pa, torch_cuda):
    torch = torch_cuda
    n, seed, lo, hi = 1000, 12345, -1.0, 1.0
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    h1pa.Randomize(x, seed, lo, hi)
    got = x.cpu().numpy()
    want = np.array([lo + (hi - lo) * ((splitmix64(seed ^ splitmix64(i)) >> 11) * 2.0 ** -53)
                     for i in range(n)])
    assert np.array_equal(got, want)


def _serve(*args):
    print(f"[serve] {' '.join(args)}", flush=True)
    r = subprocess.run([SERVE, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("mode,prec", [("batch", "bf16"), ("batch", "fp64"), ("serial", "fp64"),
                                       ("mg", "fp64")])
def test_serve_small_mix(mode, prec, torch_cuda):
    out = _serve("--mode", mode, "--precision", prec, "--rate", "60", "--seconds", "0.3",
                 "--nmin", "3", "--nmax", "8", "--pmin", "1", "--pmax", "4", "--seed", "5")
    print(out, flush=True)
    assert out["requests"] > 3
    assert out["converged"] == out["requests"]
    assert out["solves_per_s"] > 0
    assert 0 < out["latency_p50_s"] <= out["latency_p99_s"] <= out["latency_max_s"]


def test_serve_budget_limits_concurrency(torch_cuda):
    """A budget that fits one request at a time still serves every request."""
    out = _serve("--mode", "batch", "--rate", "200", "--seconds", "0.1", "--nmin", "4",
                 "--nmax", "6", "--pmin", "2", "--pmax", "3", "--budget-gb", "0.00001",
                 "--seed", "9")
    assert out["converged"] == out["requests"]
    assert out["peak_live"] == 1


def _rhs(seed, n):
    """h1pa_vector_randomize(seed, -1, 1) on the host (the generator is pinned
    bit for bit by test_vector_randomize_bit_exact)."""
    with np.errstate(over="ignore"):
        def sm(x):
            x = x + np.uint64(0x9E3779B97F4A7C15)
            x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return x ^ (x >> np.uint64(31))
        i = np.arange(n, dtype=np.uint64)
        u = (sm(np.uint64(seed) ^ sm(i)) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return -1.0 + 2.0 * u


def _read(d):
    reqs = {}
    for line in open(os.path.join(d, "requests.txt")):
        i, n, p, cs, rs, it, conv = line.split()
        reqs[int(i)] = dict(n=int(n), p=int(p), cseed=int(cs), rseed=int(rs), it=int(it),
                            conv=int(conv))
    return reqs


def test_serve_cfg5_mix_equals_alone(orc, torch_cuda, tmp_path):
    """Continuous batching on a draw of the serving workload (n in [6, 32]
    with P(n) ~ 1/n, p in 1..8, BF16-PA mixed precision at 1e-10): every
    served solution equals, bit for bit, the same request solved alone
    through the batch (--solo), with the same iteration count; and every
    request up to 4M DOFs meets CGSolver's measure of the FP64 true residual,
    computed by the oracle from the request's own mesh, coefficient and RHS."""
    mix, solo = tmp_path / "mix", tmp_path / "solo"
    mix.mkdir()
    solo.mkdir()
    common = ["--mode", "batch", "--precision", "bf16", "--rate", "200", "--seconds", "0.1",
              "--seed", "11"]
    out_mix = _serve(*common, "--dump", str(mix))
    out_solo = _serve(*common, "--solo", "--dump", str(solo))
    print(out_mix, out_solo, flush=True)
    assert out_solo["peak_live"] == 1 and out_mix["peak_live"] > 1
    rm, rs = _read(mix), _read(solo)
    assert rm.keys() == rs.keys() and len(rm) >= 8
    for i, q in rm.items():
        assert q["conv"] == 1 and rs[i]["conv"] == 1
        assert (q["n"], q["p"], q["cseed"], q["rseed"]) == \
            (rs[i]["n"], rs[i]["p"], rs[i]["cseed"], rs[i]["rseed"])
        assert q["it"] == rs[i]["it"], i
        a = (mix / f"x_{i}.bin").read_bytes()
        assert a == (solo / f"x_{i}.bin").read_bytes(), f"request {i} differs alone"
    checked = 0
    for i, q in sorted(rm.items()):
        n, p = q["n"], q["p"]
        if (n * p + 1) ** 3 > 4_000_000:
            continue
        P = orc.Problem(n, p, perturb=0.15, seed=q["cseed"], coeff=("fourier", q["cseed"]))
        b = _rhs(q["rseed"], P.ndofs)
        _, Bo = P.form_linear_system(np.zeros(P.ndofs), b)
        x = np.fromfile(mix / f"x_{i}.bin", dtype=np.float64)
        r = Bo - P.constrained_mult(x)
        dinv = 1.0 / P.diagonal()
        dinv[P.ess] = 1.0
        meas = np.sqrt(r @ (dinv * r)) / np.sqrt(Bo @ (dinv * Bo))
        print(f"[serve] request {i} n={n} p={p} it={q['it']} measure {meas:.2e}", flush=True)
        assert meas <= 1.05e-10
        checked += 1
    assert checked >= 5


def test_serve_mg_concurrent_equals_one_worker(orc, torch_cuda, tmp_path):
    """Multigrid serving on a draw of the serving workload: the results of 4
    solver threads running requests side by side equal, bit for bit, those of
    one solver thread, with the same iteration counts; every result meets the
    serial mode's stop measure (Jacobi-weighted FP64 true residual <= 1e-10,
    checked by the driver on every request and by the oracle on those up to
    4M DOFs)."""
    four, one = tmp_path / "four", tmp_path / "one"
    four.mkdir()
    one.mkdir()
    common = ["--mode", "mg", "--precision", "fp64", "--rate", "200", "--seconds", "0.1",
              "--seed", "11"]
    out4 = _serve(*common, "--workers", "4", "--dump", str(four))
    out1 = _serve(*common, "--workers", "1", "--dump", str(one))
    print(out4, out1, flush=True)
    assert out4["converged"] == out4["requests"] == out1["requests"]
    assert out4["max_jacobi_rel_residual"] <= 1.05e-10
    r4, r1 = _read(four), _read(one)
    assert r4.keys() == r1.keys() and len(r4) >= 8
    for i, q in r4.items():
        assert q["conv"] == 1 and q["it"] == r1[i]["it"], i
        assert (four / f"x_{i}.bin").read_bytes() == (one / f"x_{i}.bin").read_bytes(), i
    checked = 0
    for i, q in sorted(r4.items()):
        n, p = q["n"], q["p"]
        if (n * p + 1) ** 3 > 4_000_000:
            continue
        P = orc.Problem(n, p, perturb=0.15, seed=q["cseed"], coeff=("fourier", q["cseed"]))
        b = _rhs(q["rseed"], P.ndofs)
        _, Bo = P.form_linear_system(np.zeros(P.ndofs), b)
        x = np.fromfile(four / f"x_{i}.bin", dtype=np.float64)
        r = Bo - P.constrained_mult(x)
        dinv = 1.0 / P.diagonal()
        dinv[P.ess] = 1.0
        meas = np.sqrt(r @ (dinv * r)) / np.sqrt(Bo @ (dinv * Bo))
        print(f"[serve-mg] request {i} n={n} p={p} it={q['it']} measure {meas:.2e}", flush=True)
        assert meas <= 1.05e-10
        checked += 1
    assert checked >= 5
