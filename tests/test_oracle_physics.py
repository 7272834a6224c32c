"""Pins the oracle's non-FFT functions with closed forms, library routines
and invariants (no reference run needed)."""
import math

import numpy as np
import pytest
from scipy.special import erfc

import oracle_lib as O


def setup32():
    return O.Setup(1, 20.0, 32, [0, 0], [4.0])


def test_nwave_matches_reference_aimd_h():
    # Example0-double-serial/aimd.h:1 holds NWAVE = 8932 for the 32^3, 20 bohr cell
    s = setup32()
    assert int(s.masker[0::2].sum()) == 8932


def test_lattice_and_tables():
    s = O.Setup(1, 20.0, 16, [0], [1.0])
    assert s.omega == pytest.approx(8000.0, rel=1e-15)
    b = 2 * math.pi / 20.0
    assert np.allclose(s.unitg.reshape(3, 3), b * np.eye(3), rtol=0, atol=1e-15)
    G = s.G.reshape(3, -1)
    gg = (G ** 2).sum(0)
    m = s.masker[0::2].astype(bool)
    # mask: strict |G|^2 < (b N/2)^2 and kx < N/2, |ky|,|kz| < N/2
    n, nh = 16, 9
    kx = np.arange(s.nfft3d) % nh
    ky = (np.arange(s.nfft3d) // nh) % n
    kz = np.arange(s.nfft3d) // (nh * n)
    inside = (gg < (b * 8) ** 2) & (kx < 8) & (ky != 8) & (kz != 8)
    assert np.array_equal(m, inside)
    assert np.allclose(s.tg[m], 0.5 * gg[m]) and np.all(s.tg[~m] == 0)
    nz = m & (gg > 0)
    assert np.allclose(s.vc[nz], 4 * math.pi / gg[nz], rtol=1e-15)
    w = 0.25 * s.rcut ** 2
    assert np.allclose(s.vg[nz], 4 * math.pi / gg[nz] * np.exp(-w * gg[nz]), rtol=1e-14)
    assert s.vc[0] == 0 and s.vg[0] == 0


def test_gcdotc_equals_full_sphere_and_parseval():
    n = 12
    s = O.Setup(1, 10.0, n, [0], [1.0])
    rng = np.random.default_rng(3)
    # two real fields -> half-complex coefficients (Hermitian kx=0 plane)
    fa, fb = rng.standard_normal((2, n, n, n))
    a = np.zeros((n, n, n + 2)); a[:, :, :n] = fa
    b = np.zeros((n, n, n + 2)); b[:, :, :n] = fb
    A = O.rc_fft3f(a.reshape(-1), n)
    B = O.rc_fft3f(b.reshape(-1), n)
    # zero the kx = N/2 column, which gcdotc would count twice for a real field
    A = A.reshape(n, n, n + 2); B = B.reshape(n, n, n + 2)
    A[:, :, n:] = 0; B[:, :, n:] = 0
    g = O.gcdotc(n, A.reshape(-1), B.reshape(-1))
    # full-sphere sum Re sum conj(A) B over all G of the Hermitian extension
    Xa = np.fft.fftn(fa); Xb = np.fft.fftn(fb)
    Xa[:, :, n // 2] = 0; Xb[:, :, n // 2] = 0
    full = np.sum((np.conj(Xa) * Xb).real)
    assert g == pytest.approx(full, rel=1e-12)


def test_phafac_strfac_vs_direct_exponential():
    n = 16
    s = O.Setup(1, 12.0, n, [0, 0], [1.0, 1.0])
    r = np.array([0.7, -1.3, 2.9, -3.1, 0.4, 1.1])
    ex1 = np.zeros(2 * (n + 2)); ex2 = np.zeros(2 * 2 * n); ex3 = np.zeros(2 * 2 * n)
    O.lib().oracle_phafac(n, s.unitg, 2, r, ex1, ex2, ex3)
    G = s.G.reshape(3, -1)
    nh = n // 2 + 1
    kx = np.arange(s.nfft3d) % nh
    ky = (np.arange(s.nfft3d) // nh) % n
    kz = np.arange(s.nfft3d) // (nh * n)
    sk = lambda k: np.where(k > n // 2, k - n, k)
    nyq = (kx == n // 2) | (ky == n // 2) | (kz == n // 2)
    for ii in range(2):
        exi = np.zeros(2 * s.nfft3d)
        O.lib().oracle_strfac(n, ex1[ii * (n + 2):], ex2[ii * 2 * n:], ex3[ii * 2 * n:], exi)
        got = exi[0::2] + 1j * exi[1::2]
        ref = np.exp(-1j * (G.T @ r[3 * ii:3 * ii + 3])) * (-1.0) ** (kx + sk(ky) + sk(kz))
        ref[nyq] = 0
        assert np.max(np.abs(got - ref)) < 1e-13


def test_slater_exchange_textbook():
    rho = np.array([1e-6, 1e-3, 0.01, 0.1, 1.0, 7.5])  # per-spin density, ispin=1
    vx = np.zeros_like(rho); ex = np.zeros_like(rho)
    O.lib().oracle_dirac_exchange(len(rho), rho, vx, ex)
    n = 2 * rho
    ref_ex = -(3.0 / 4.0) * (3.0 / math.pi) ** (1.0 / 3.0) * n ** (1.0 / 3.0)
    assert np.allclose(ex, ref_ex, rtol=1e-14)
    assert np.allclose(vx, 4.0 / 3.0 * ref_ex, rtol=1e-14)


def vwn5_para(rs):
    """Vosko-Wilk-Nusair (1980) paramagnetic fit 'V' with its published
    parameters A = 0.0310907, x0 = -0.10498, b = 3.72744, c = 12.9352."""
    A, x0, b, c = 0.0310907, -0.10498, 3.72744, 12.9352
    x = math.sqrt(rs)
    X = lambda t: t * t + b * t + c
    Q = math.sqrt(4 * c - b * b)
    return A * (math.log(x * x / X(x)) + 2 * b / Q * math.atan(Q / (2 * x + b))
                - b * x0 / X(x0) * (math.log((x - x0) ** 2 / X(x)) + 2 * (b + 2 * x0) / Q
                                    * math.atan(Q / (2 * x + b))))


@pytest.mark.parametrize("rs", [0.5, 1.0, 2.0, 5.0, 10.0, 30.0])
def test_vwn_energy_vs_published_formula(rs):
    # rho per spin such that rs = (3/(4 pi n))^(1/3), n = 2 rho
    n = 3.0 / (4.0 * math.pi * rs ** 3)
    rho = np.array([n / 2])
    xcp = np.zeros(1); xce = np.zeros(1)
    O.lib().oracle_vosko(1, rho, xcp, xce)
    # the reference's derived constants are 7-digit literals: agree to ~1e-6
    assert xce[0] == pytest.approx(vwn5_para(rs), rel=2e-6, abs=1e-9)


@pytest.mark.parametrize("rs", [0.7, 1.5, 4.0, 12.0])
def test_vwn_potential_is_derivative(rs):
    """v_c = d(n e_c)/dn = e_c - (rs/3) de_c/drs, by numerical differentiation
    of the oracle's own energy."""
    def ec(r):
        n = 3.0 / (4.0 * math.pi * r ** 3)
        xcp = np.zeros(1); xce = np.zeros(1)
        O.lib().oracle_vosko(1, np.array([n / 2]), xcp, xce)
        return xce[0], xcp[0]
    h = 1e-5 * rs
    e0, v0 = ec(rs)
    d = (ec(rs + h)[0] - ec(rs - h)[0]) / (2 * h)
    assert v0 == pytest.approx(e0 - rs / 3 * d, rel=2e-6)


def test_vosko_zero_density_is_finite():
    xcp = np.zeros(3); xce = np.zeros(3)
    O.lib().oracle_vosko(3, np.zeros(3), xcp, xce)
    assert np.all(np.isfinite(xcp)) and np.all(np.isfinite(xce))


def test_erfc_as_polynomial():
    w = np.linspace(0, 6, 301)
    got = np.array([O.lib().oracle_erfc_as(x) for x in w])
    # Abramowitz & Stegun 7.1.28: |error| <= 3e-7
    assert np.max(np.abs(got - erfc(w))) < 3e-7
    assert got[0] == 1.0


def test_single_ion_ewald_is_madelung():
    """One ion in a simple cubic cell: the G sum cancels against cewald,
    leaving -z^2 alpha / (2 rs) with the Madelung constant of main.c:150."""
    s = O.Setup(1, 15.0, 24, [0], [3.0])
    # eion = 1/2 sum |S|^2 vg / omega + cewald, |S(G)| = z on the mask
    vt = (3.0 ** 2) * s.vg
    eion = O.lib().oracle_gssum(24, vt) * 0.5 / s.omega + s.cewald
    rs = (3 * s.omega / (4 * math.pi)) ** (1 / 3)
    assert eion == pytest.approx(-9.0 * 1.76011888 / (2 * rs), rel=1e-12)


def test_simpson_exact_for_cubics():
    h = 0.02
    x = np.arange(501) * h
    y = 1.5 - 2 * x + 0.7 * x ** 2 + 0.3 * x ** 3
    F = lambda t: 1.5 * t - t ** 2 + 0.7 / 3 * t ** 3 + 0.3 / 4 * t ** 4
    assert O.lib().oracle_simp(501, y, h) == pytest.approx(F(x[-1]) - F(0), rel=1e-13)


def test_syev_sorted():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((7, 7)); a = a + a.T
    A = np.asfortranarray(a).reshape(-1, order="F").copy()
    e = np.zeros(7)
    O.lib().oracle_syev_sorted(7, A, e)
    V = A.reshape(7, 7, order="F")
    assert np.allclose(e, np.sort(np.linalg.eigvalsh(a))[::-1], atol=1e-12)
    assert np.allclose(a @ V, V * e, atol=1e-11)
