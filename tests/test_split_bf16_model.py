"""The arithmetic precision mode 3 rests on, checked on the CPU with numpy.

- The split of an FP32 value x into BF16 halves, hi = bf16(x) and
  lo = bf16(x - hi), both rounded to nearest even by the integer formula of
  the device (bf16_rne: (u + 0x7fff + ((u >> 16) & 1)) >> 16 on the FP32
  bits), leaves |x - hi - lo| <= 2^-17 |x|, and hi is numpy's own RNE of x
  to 8 mantissa bits.
- A product from three BF16 products, hi hi + hi lo + lo hi, is within
  2^-15 of x y relative; a dot product of 4096 such terms summed in FP32 is
  within 2^-18 of the exact one relative to the sum of |terms| (2.5e-7
  measured; the high halves alone miss by 1.6e-4, so the bound catches a
  dropped low half).
- The overlaps of the step d = c2 - c1:
      <c2|c1> = <c1|c1> + <d|c1>,
      <c2|c2> = <c1|c1> + <d|c1> + <c1|d> + <d|d>
  hold to FP64 rounding, and with the products of d split as above the
  overlaps are off by a few 1e-7 |d| (the FP32 sums), not by 2^-16 |c|.
"""
import numpy as np

SPLIT_TERM = 2.0 ** -17


def bf16_rne_bits(x):
    """the device's round-to-nearest-even BF16 bits of finite FP32 values"""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def split(x):
    x = np.asarray(x, dtype=np.float32)
    hi = bf16_value(bf16_rne_bits(x))
    lo = bf16_value(bf16_rne_bits((x - hi).astype(np.float32)))
    return hi, lo


def split_dot(a, b, w=None):
    """sum_k w_k a_k b_k as the split-BF16 kernels form it: FP64 inputs
    rounded to FP32, split, three products per term, FP32 sums"""
    a32 = np.asarray(a, dtype=np.float64).astype(np.float32)
    b64 = np.asarray(b, dtype=np.float64) * (1.0 if w is None else w)
    b32 = b64.astype(np.float32)
    ah, al = split(a32)
    bh, bl = split(b32)
    terms = (ah.astype(np.float32) * bh + ah * bl + al * bh).astype(np.float32)
    return np.sum(terms, dtype=np.float32).astype(np.float64)


def test_rne_formula_matches_numpy_rounding():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(200000) * np.exp(rng.uniform(-30, 30, 200000))).astype(np.float32)
    hi = bf16_value(bf16_rne_bits(x))
    # round to 8 mantissa bits, nearest even, by numpy: scale to the ulp
    m, e = np.frexp(x.astype(np.float64))
    ref = np.ldexp(np.round(m * 256.0), e - 8)   # numpy rounds half to even
    assert np.array_equal(hi.astype(np.float64), ref)


def test_split_leaves_a_2_to_minus_17_remainder():
    rng = np.random.default_rng(2)
    x = (rng.standard_normal(200000) * np.exp(rng.uniform(-20, 20, 200000))).astype(np.float32)
    hi, lo = split(x)
    rem = np.abs(x.astype(np.float64) - hi.astype(np.float64) - lo.astype(np.float64))
    assert np.all(rem <= SPLIT_TERM * np.abs(x.astype(np.float64)) * 1.0000001)
    # the remainder of the high half alone is up to 2^-9 |x|: the low half matters
    assert np.max(np.abs(x - hi) / np.abs(x)) > 2.0 ** -10


def test_three_products_within_2_to_minus_15():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(100000).astype(np.float32)
    y = rng.standard_normal(100000).astype(np.float32)
    xh, xl = split(x)
    yh, yl = split(y)
    p = xh.astype(np.float64) * yh + xh.astype(np.float64) * yl + xl.astype(np.float64) * yh
    exact = x.astype(np.float64) * y.astype(np.float64)
    assert np.all(np.abs(p - exact) <= 2.0 ** -15 * np.abs(exact) + 1e-300)


def test_dot_product_error_bound():
    rng = np.random.default_rng(4)
    n = 4096
    for _ in range(20):
        a = rng.standard_normal(n)
        b = rng.standard_normal(n)
        w = np.where(rng.random(n) < 0.02, 1.0, 2.0)
        exact = np.sum(w * a * b)
        got = split_dot(a, b, w)
        scale = np.sum(np.abs(w * a * b))
        # FP32 accumulation of n terms plus the split
        assert abs(got - exact) <= 2.0 ** -18 * scale


def test_step_overlaps_from_d():
    """the assembly of k_gram_dfix: exact to FP64 rounding, and the split
    products' error scales with |d|"""
    rng = np.random.default_rng(5)
    ne, n = 8, 3000
    c1 = np.linalg.qr(rng.standard_normal((n, ne)))[0].T          # orthonormal rows
    for dscale in (1e-2, 1e-4, 1e-6):
        d = dscale * rng.standard_normal((ne, n))
        c2 = c1 + d
        s11 = c1 @ c1.T
        A = d @ c1.T                       # <d_i|c1_j>
        B = d @ d.T                        # <d_i|d_j>
        g21 = s11 + A
        g22 = s11 + (A + A.T) + B
        assert np.max(np.abs(g21 - c2 @ c1.T)) < 1e-13
        assert np.max(np.abs(g22 - c2 @ c2.T)) < 1e-13
        # the same with A and B from split-BF16 dot products
        As = np.array([[split_dot(d[i], c1[j]) for j in range(ne)] for i in range(ne)])
        Bs = np.array([[split_dot(d[i], d[j]) for j in range(ne)] for i in range(ne)])
        err21 = np.max(np.abs(s11 + As - c2 @ c1.T))
        err22 = np.max(np.abs(s11 + (As + As.T) + Bs - c2 @ c2.T))
        dn = np.max(np.linalg.norm(d, axis=1))
        # |<d|c1>| <= |d|: the error is a small multiple of 1e-7 |d|
        # (measured 2e-7 and 2-17e-7; the high halves alone give 1e-4)
        assert err21 <= 2.0 ** -18 * dn
        assert err22 <= 2.0 ** -16 * dn
