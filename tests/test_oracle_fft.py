"""Pins the oracle's FFT layer (the restatement of fftpack3d/fftpack3d.f and
the FFTPACK conventions) against brute-force DFTs and numpy's independent
FFT implementation."""
import numpy as np
import pytest

import oracle_lib as O


def brute_dft(x, sign):
    n = len(x)
    k = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 20, 24, 25, 30, 32, 48, 96, 128])
@pytest.mark.parametrize("sign", [-1, 1])
def test_cfft1d_vs_brute(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = np.empty(2 * n)
    a[0::2], a[1::2] = x.real, x.imag
    O.lib().oracle_cfft1d(n, sign, a)
    got = a[0::2] + 1j * a[1::2]
    ref = brute_dft(x, sign)
    assert np.max(np.abs(got - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


def full_grid_rfft(r):
    """reference-layout forward transform of a real [z][y][x] field via numpy"""
    n = r.shape[0]
    X = np.fft.rfftn(r, axes=(0, 1, 2))  # x last: (n, n, n/2+1)
    out = np.zeros((n, n, n + 2))
    out[:, :, 0::2] = X.real
    out[:, :, 1::2] = X.imag
    return out


@pytest.mark.parametrize("n", [6, 8, 10, 12, 16, 24, 32])
def test_rc_fft3f_vs_numpy(n):
    rng = np.random.default_rng(100 + n)
    r = rng.standard_normal((n, n, n))
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = r
    a[:, :, n:] = rng.standard_normal((n, n, 2))  # pads are ignored
    got = O.rc_fft3f(a.reshape(-1), n).reshape(n, n, n + 2)
    ref = full_grid_rfft(r)
    assert np.max(np.abs(got - ref)) < 1e-12 * n ** 3
    # Im(kx=0) and Im(kx=N/2) are exactly zero after the x pass
    # (fftpack3d.f:160-161) -- before the y/z passes mix them in.
    a2 = np.zeros((1, 1, n + 2))


@pytest.mark.parametrize("n", [6, 8, 12, 16, 24])
def test_cr_fft3b_brute_force(n):
    """cr_fft3b on arbitrary half-complex input (non-Hermitian kx=0 plane
    included) against the literal definition: z, y complex backward passes,
    then drfftb on each x row using Re X_0, X_1..X_{N/2-1}, Re X_{N/2}."""
    rng = np.random.default_rng(7 * n)
    nh = n // 2 + 1
    A = rng.standard_normal((n, n, nh)) + 1j * rng.standard_normal((n, n, nh))  # [kz][ky][kx]
    k = np.arange(n)
    W = np.exp(2j * np.pi * np.outer(k, k) / n)
    B = np.einsum("zk,kyx->zyx", W, A)        # z pass
    B = np.einsum("yk,zkx->zyx", W, B)        # y pass
    x = np.arange(n)
    out = np.zeros((n, n, n))
    for q in range(nh):
        if q == 0:
            term = B[:, :, 0].real[:, :, None] * np.ones(n)
        elif q == n // 2:
            term = B[:, :, q].real[:, :, None] * ((-1.0) ** x)
        else:
            ph = np.exp(2j * np.pi * q * x / n)
            term = 2.0 * (B[:, :, q][:, :, None] * ph).real
        out += term
    a = np.zeros((n, n, nh, 2))
    a[..., 0], a[..., 1] = A.real, A.imag
    got = O.cr_fft3b(a.reshape(-1), n).reshape(n, n, n + 2)
    assert np.max(np.abs(got[:, :, :n] - out)) < 1e-11 * n ** 3
    assert np.all(got[:, :, n:] == 0.0)


@pytest.mark.parametrize("n", [8, 12, 16, 32])
def test_roundtrip_scales_by_n3(n):
    rng = np.random.default_rng(n)
    r = rng.standard_normal((n, n, n))
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = r
    back = O.cr_fft3b(O.rc_fft3f(a.reshape(-1), n), n).reshape(n, n, n + 2)
    assert np.max(np.abs(back[:, :, :n] - n ** 3 * r)) < 1e-10 * n ** 3
