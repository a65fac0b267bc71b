"""The oracle's BF16 storage rule against torch's float32 -> bfloat16
conversion (round to nearest even), including ties, signs, zeros, the
largest finite values and NaN/inf."""
import numpy as np


def test_round_bf16_matches_torch(orc):
    import torch
    rng = np.random.default_rng(0)
    a = np.concatenate([
        rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000),
        np.array([0.0, -0.0, 1.0, -1.0, 3.0e38, -3.0e38, np.inf, -np.inf]),
    ])
    # exact ties in float32: low 16 bits == 0x8000, with even and odd kept bit
    # (kept bits: sign, exponent 0x3F..0x41 range, 7 mantissa bits)
    kept = (rng.integers(0, 1 << 7, 1000, dtype=np.uint32) << 16) | np.uint32(0x3F800000)
    kept |= rng.integers(0, 2, 1000, dtype=np.uint32) << 31
    ties = (kept | 0x8000).view(np.float32).astype(np.float64)
    odd_ties = (kept | 0x10000 | 0x8000).view(np.float32).astype(np.float64)
    a = np.concatenate([a, ties, odd_ties])
    want = torch.from_numpy(a.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    got = orc.round_bf16(a)
    assert np.array_equal(got, want)
    assert np.isnan(orc.round_bf16(np.array([np.nan]))[0])


def test_round_bf16_error_bound(orc):
    """|round(x) - x| <= 2^-8 |x| for normal values: 8 significant bits, so
    half an ulp is at most 2^-8 of the value; and some values reach 2^-9."""
    a = np.random.default_rng(1).uniform(-1e3, 1e3, 100000)
    r = orc.round_bf16(a)
    f32 = a.astype(np.float32).astype(np.float64)
    rel = np.abs(r - f32) / np.abs(f32)
    assert np.all(rel <= 2.0 ** -8)
    assert rel.max() > 2.0 ** -9
