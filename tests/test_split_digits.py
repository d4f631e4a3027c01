"""The split form's digit bounds (steppe_amd/csrc/f2_split.hip, prep_split_kernel),
restated in numpy from the kernel's own steps and checked against what the scheme needs:

* q' = q - 1/2 as six balanced 8-bit digits: |d_0| <= 64, |d_k| <= 128 (exact in BF16);
* w = -(q~^2 - h)/2 as four FP16 digits A, 32 b, 4 c, D on the levels with units 2^-14,
  2^-30, 2^-38, 2^-46: each exactly representable in FP16 and |w - w~| <= 2^-47
  (checked exactly with fractions on the worst cases);
* per SNP, the growth of every FP32 level accumulator keeps 128 SNPs below 2^24.

Counts run over every population size of the panels (n = 1 .. 16383) with edge cases
(q = 0, 1/2, 1 and q = 1/256-style values with short binary expansions).
"""
from fractions import Fraction

import numpy as np


def _digits(a, n):
    q = a / (2.0 * n)
    h = q * (1.0 - q) / np.maximum(n - 1.0, 1.0)
    x = (q - 0.5) * 128.0
    d = []
    for _ in range(6):
        dk = np.rint(x)
        d.append(dk)
        x = (x - dk) * 256.0
    qt = (d[0] + (d[1] + (d[2] + (d[3] + (d[4] + d[5] / 256) / 256) / 256) / 256) / 256) / 128
    w = -0.5 * (qt * qt - h)
    y = w * 16384.0
    A = np.rint(y)
    r1 = (y - A) * 2048.0
    b = np.rint(r1)
    r2 = (r1 - b) * 2048.0
    c = np.rint(r2)
    D = np.rint((r2 - c) * 1024.0)
    return d, w, A, 32.0 * b, 4.0 * c, D


def _counts():
    rng = np.random.default_rng(7)
    n = rng.integers(1, 16384, 2_000_000).astype(np.float64)
    a = np.minimum(np.floor(rng.random(n.size) * (2 * n + 1)), 2 * n)
    edge_n = np.array([1, 1, 1, 2, 2, 128, 128, 16383, 16383, 16383, 256, 512], np.float64)
    edge_a = np.array([0, 1, 2, 1, 3, 1, 255, 0, 16383, 32766, 1, 3], np.float64)
    return np.concatenate([edge_a, a]), np.concatenate([edge_n, n])


def test_q_digits_fit_bf16():
    a, n = _counts()
    d, *_ = _digits(a, n)
    assert np.abs(d[0]).max() <= 64
    for dk in d[1:]:
        assert np.abs(dk).max() <= 128
    # integers with at most 8 significant bits are exact in BF16 (and d5 / 256 too)
    for dk in d:
        f = dk.astype(np.float32).view(np.uint32)
        assert np.all(f & 0xFFFF == 0)


def test_w_digits_fit_fp16_and_reconstruct_w():
    a, n = _counts()
    _, w, A, B, C, D = _digits(a, n)
    assert np.abs(A).max() <= 2048 and np.abs(B).max() <= 32 * 1024
    assert np.abs(C).max() <= 4 * 1024 and np.abs(D).max() <= 512
    for v in (A, B, C, D):
        assert np.array_equal(v.astype(np.float16).astype(np.float64), v)
    wt = A * 2.0 ** -14 + B * 2.0 ** -30 + C * 2.0 ** -38 + D * 2.0 ** -46
    err = np.abs(w - wt)
    assert err.max() <= 2.0 ** -47 * (1 + 1e-9)
    for i in np.argsort(err)[-20:]:
        exact = (Fraction(int(A[i]), 2 ** 14) + Fraction(int(B[i]), 2 ** 30) + Fraction(int(C[i]), 2 ** 38)
                 + Fraction(int(D[i]), 2 ** 46))
        assert abs(Fraction(float(w[i])) - exact) <= Fraction(1, 2 ** 47)


def test_level_accumulators_stay_exact_over_128_snps():
    a, n = _counts()
    d, _, A, B, C, D = _digits(a, n)
    m = [np.abs(x).max() for x in d]
    # worst case per SNP: the largest digits on both sides of every product of a level
    levels = [
        m[0] * m[0] + 2 * np.abs(A).max(),
        2 * m[0] * m[1],
        2 * m[0] * m[2] + m[1] * m[1] + 2 * np.abs(B).max(),
        2 * m[0] * m[3] + 2 * m[1] * m[2] + 2 * np.abs(C).max(),
        2 * m[0] * m[4] + 2 * m[1] * m[3] + m[2] * m[2]
        + (2 * m[0] * m[5] + 2 * m[1] * m[4] + 2 * m[2] * m[3]) / 256 + 2 * np.abs(D).max(),
    ]
    for t, g in enumerate(levels):
        # 128 SNPs plus the carry a level receives from the one below it (< 2^16)
        assert 128 * g + 2 ** 16 < 2 ** 24, (t, g)
