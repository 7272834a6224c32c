"""The oracle's preconditioned step (oracle_set_preconditioner), which the
GPU path is checked against: with a per-G factor that varies strongly across
G, the step's component along c1 is taken out of the overlaps and added back
through lmd = X - D/dte.  Pinned by what the constraint must deliver whatever
K is: orthonormal orbitals after every call (a wrong sign or transpose in the
projected overlaps leaves the Lagrange fixed point solving the wrong
equation), an energy that decreases, and, with K = 0 on every G, no change
at all beyond the constraint's rotation within the occupied space."""
import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O


def tpa(tg, kmax):
    x = tg
    p = 27.0 + 18.0 * x + 12.0 * x ** 2 + 8.0 * x ** 3
    return kmax * p / (p + 16.0 * x ** 4)


def setup(ne=12):
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    st = O.State(O.Setup(1, 12.0, 24, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl, cell.vnlnrm,
                 cell.r, psi, cell.amass)
    return cell, st


def overlap(cell, st, ne):
    c2 = st.c2.reshape(ne, -1)
    w = np.full(cell.n2ft3d, 2.0)
    w.reshape(24 * 24, 26)[:, 0:2] = 1.0
    return (c2 * w) @ c2.T


@pytest.mark.parametrize("kmax", [4.0, 16.0])
def test_precond_calls_stay_orthonormal_and_descend(kmax):
    ne = 12
    cell, st = setup(ne)
    K = tpa(cell.tg, kmax)
    h, Es = None, []
    for call in range(5):
        E, _, _ = st.step(5, K=None if h is None else K, h=h)
        h = st.hml.copy()
        Es.append(E[0])
        S = overlap(cell, st, ne)
        assert np.max(np.abs(S - np.eye(ne))) < 1e-12, call
    assert np.all(np.diff(Es) < 0)


def test_zero_factor_keeps_the_occupied_space():
    """K = 0: c2 = c1 before the overlaps, so D = <c1|c1> - I = 0 and the
    constraint can only rotate within span(c1); the projector onto the
    occupied space is unchanged by a call"""
    ne = 12
    cell, st = setup(ne)
    st.step(3)                          # a plain call gives h
    h = st.hml.copy()
    w = np.full(cell.n2ft3d, 2.0)
    w.reshape(24 * 24, 26)[:, 0:2] = 1.0
    c_before = st.c2.reshape(ne, -1).copy()
    st.step(1, K=np.zeros(cell.nfft3d), h=h)
    c_after = st.c2.reshape(ne, -1)
    # components of the new orbitals outside span(c_before)
    proj = (c_after * w) @ c_before.T
    resid = c_after - proj @ c_before
    assert np.max(np.abs(resid)) < 1e-12
