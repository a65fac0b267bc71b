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


@pytest.mark.parametrize("mode,prec", [("batch", "bf16"), ("batch", "fp64"), ("serial", "fp64")])
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
