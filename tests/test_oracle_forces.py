"""Ionic forces and the ion update of the oracle (move = 1,
ccpsdv3/cpsd.c:246-257, 291-304, 310-364, 401-411).

No reference run of move = 1 is recorded, so the forces are pinned by the
mathematics: at fixed orbitals the reference's local, nonlocal and Ewald
forces are -dE0/dR of its own total energy E0 (pinned separately against the
Example0 golden energy).  A central difference of E0 over each ion coordinate
must reproduce every force component; a dropped term, a wrong sign or a
missing factor 2 fails it.  The one deliberate inconsistency of the
reference, the Abramowitz-Stegun erfc whose derivative is taken as the true
erfc's (2/sqrt(pi)) w exp(-w^2), is added in closed form.
"""
import copy

import numpy as np
import pytest

import oracle_lib as O
from example_inputs import example0_oracle_state, water_oracle_state


def energy_at(st, r):
    """E0 at positions r with the state's current orbitals (one step from c2)."""
    t = copy.deepcopy(st)
    t.r2[:] = r
    t.E[:] = 0.0
    E, _, _ = t.step(1, move=0)
    return E[0]


B = (0.0705230784, 0.0422820123, 0.0092705272, 0.0001520143, 0.0002765672, 0.0000430638)
CERFC = 1.128379167


def erfc_as_deriv(w):
    """d/dw of (1 + sum B_k w^k)^-16, the erfc of ccpsdv3/cpsd.c:11-17."""
    p = 1.0 + sum(b * w ** (k + 1) for k, b in enumerate(B))
    dp = sum((k + 1) * b * w ** k for k, b in enumerate(B))
    return -16.0 * dp * p ** -17


def ewald_derivative_mismatch(st, r):
    """F_reference - (-dE0/dR): the real-space Ewald force takes the true erfc
    derivative -CERFC exp(-w^2) where the energy holds the A&S polynomial, so
    the two differ by zz/(rcut r) (erfc_AS'(w) + CERFC exp(-w^2)) x/r per pair
    and image."""
    s = st.s
    nion, nsh = s.nion, 27
    rc = s.rcell.reshape(3, nsh)   # [x | y | z] of the nsh image shifts
    zv = s.zv[s.katm]
    R = r.reshape(-1, 3)
    d = np.zeros_like(R)
    for i in range(nion):
        for j in range(i + 1, nion):
            for l in range(nsh):
                x = rc[:, l] + R[i] - R[j]
                rr = np.sqrt(x @ x)
                w = rr / s.rcut
                g = zv[i] * zv[j] / (s.rcut * rr) * (erfc_as_deriv(w) + CERFC * np.exp(-w * w)) * x / rr
                d[i] += g
                d[j] -= g
    return d.reshape(-1)


def forces_at(st):
    t = copy.deepcopy(st)
    t.step(1, move=1)
    return t.fion.copy(), t


@pytest.mark.parametrize("name", ["example0_c2", "water"])
def test_forces_are_energy_gradient(name):
    if name == "example0_c2":
        _, st = example0_oracle_state()
    else:
        st = water_oracle_state(nfft=24, L=16.0, ne=4)
    r0 = st.r2.copy()
    f, _ = forces_at(st)
    h = 1e-4
    fd = np.zeros_like(f)
    for k in range(len(r0)):
        rp = r0.copy(); rp[k] += h
        rm = r0.copy(); rm[k] -= h
        fd[k] = -(energy_at(st, rp) - energy_at(st, rm)) / (2 * h)
    assert np.max(np.abs(f)) > 1e-3          # nontrivial forces
    corr = ewald_derivative_mismatch(st, r0)
    assert np.max(np.abs(corr)) < 1e-5
    np.testing.assert_allclose(f, fd + corr, rtol=0, atol=2e-8)


def test_ion_update_and_deltar():
    """r2 = r1 + dt/sqrt(amass) * fion (ccpsdv3/cpsd.c:401-411); deltar is the
    largest |F_I| (:675-683); fixed ions leave r2 and deltar untouched."""
    st = water_oracle_state(nfft=24, L=16.0, ne=4)
    r_start = st.r2.copy()
    t = copy.deepcopy(st)
    t.step(1, move=1)
    dti = 5.8 / np.sqrt(t.amass[t.s.katm])
    np.testing.assert_array_equal(t.r1, r_start)
    expect = t.r1 + np.repeat(dti, 3) * t.fion
    np.testing.assert_allclose(t.r2, expect, rtol=0, atol=1e-15)
    fmag = np.sqrt((t.fion.reshape(-1, 3) ** 2).sum(axis=1))
    assert t.deltar.value == pytest.approx(fmag.max(), rel=1e-15)

    u = copy.deepcopy(st)
    u.step(1, move=0)
    np.testing.assert_array_equal(u.r2, r_start)
    assert u.deltar.value == 0.0


def test_moving_ions_run_lowers_energy():
    """A few inner steps with moving ions: positions change, energies stay
    finite and the SD run still lowers E0."""
    st = water_oracle_state(nfft=24, L=16.0, ne=4)
    r_start = st.r2.copy()
    E1, _, _ = st.step(5, move=1)
    E2, _, _ = st.step(5, move=1)
    assert np.all(np.isfinite(E2))
    assert np.max(np.abs(st.r2 - r_start)) > 0.0
    assert E2[0] < E1[0]
