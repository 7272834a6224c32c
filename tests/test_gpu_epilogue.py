"""The run epilogue (ccpsdv3/main.c:451-484) on the GPU against the oracle:
the orthonormality matrix of c1, the eigenpairs of hml in eigsrt order and
the rotation c1 = c2 V, through the resident cpsd_dev_epilogue() and the
drop-in cpsd_rotate().  Eigenvectors are fixed only up to sign (and to an
orthogonal mix within equal eigenvalues), so the rotated orbitals are compared
with the oracle's up to that freedom; the rotation itself is checked exactly
against the returned V."""
import sys

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from example_inputs import example0_oracle_state, water_oracle_state
from test_gpu_parity import cell_from_state, compare

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-epilogue] " + msg, flush=True)
    sys.stdout.flush()


def clusters(eig, tol):
    """index groups of (nearly) equal eigenvalues"""
    out, cur = [], [0]
    for i in range(1, eig.size):
        if abs(eig[i] - eig[i - 1]) <= tol:
            cur.append(i)
        else:
            out.append(cur)
            cur = [i]
    out.append(cur)
    return out


def check_against_oracle(ep, ref, ne, c1_rot, c2_dev):
    np.testing.assert_allclose(ep["ortho"], ref["ortho"], rtol=0, atol=1e-10)
    scale = max(1.0, np.abs(ref["eig"]).max())
    np.testing.assert_allclose(ep["eig"], ref["eig"], rtol=0, atol=1e-10 * scale)
    assert np.all(np.diff(ep["eig"]) <= 0)
    V = ep["V"].reshape(ne, ne)          # row i = eigenvector i
    Vr = ref["V"].reshape(ne, ne)
    np.testing.assert_allclose(V @ V.T, np.eye(ne), rtol=0, atol=1e-12)
    # the rotation is c1 = c2 V with the V returned
    n2 = c2_dev.size // ne
    want = V @ c2_dev.reshape(ne, n2)
    np.testing.assert_allclose(c1_rot.reshape(ne, n2), want, rtol=0,
                               atol=1e-12 * max(1.0, np.abs(want).max()))
    # eigenvectors and rotated orbitals agree with the oracle's up to an
    # orthogonal mix inside each group of equal eigenvalues (a sign for a
    # single one): Q = V_g Vr_g^T must be orthogonal and c1_g = Q c1ref_g
    c1 = c1_rot.reshape(ne, n2)
    c1r = ref["c1"].reshape(ne, n2)
    for g in clusters(ref["eig"], 1e-7 * scale):
        Q = V[g] @ Vr[g].T
        np.testing.assert_allclose(Q @ Q.T, np.eye(len(g)), rtol=0, atol=1e-6)
        np.testing.assert_allclose(c1[g], Q @ c1r[g], rtol=0, atol=1e-8)


def test_device_epilogue_water(gpu):
    say("resident epilogue, water 32^3 ne=4")
    st = water_oracle_state()
    cell = cell_from_state(st)
    dev = cpsd_amd.DeviceSD(cell, 4)
    dev.set_orbitals(st.c2)
    for _ in range(2):
        Eg, _, _ = dev.run(5)
        Er, _, _ = st.step(5)
        compare(Eg, Er)
    before = dev.get(c2=True)
    ep = dev.epilogue()
    after = dev.get(c1=True, c2=True)
    ref = O.epilogue(4, cell.n2ft3d, cell.nfft, st.c1, st.c2, st.hml)
    check_against_oracle(ep, ref, 4, after["c1"], before["c2"])
    np.testing.assert_array_equal(after["c2"], before["c2"])      # c2 is left as it is
    np.testing.assert_allclose(ep["ortho"], np.eye(4), rtol=0, atol=1e-9)
    dev.close()


def test_device_epilogue_multi_water(gpu):
    """12 orbitals (several 16-row blocks of the Gram) on 2 waters at 24^3"""
    say("resident epilogue, two waters 24^3 ne=12")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(12, seed=3, decay=0.05)
    st = O.State(O.Setup(1, 12.0, 24, cell.katm, cell.zv), 12, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    dev = cpsd_amd.DeviceSD(cell, 12)
    dev.set_orbitals(psi)
    Eg, _, _ = dev.run(4)
    Er, _, _ = st.step(4)
    compare(Eg, Er)
    before = dev.get(c2=True)
    ep = dev.epilogue()
    after = dev.get(c1=True)
    ref = O.epilogue(12, cell.n2ft3d, cell.nfft, st.c1, st.c2, st.hml)
    check_against_oracle(ep, ref, 12, after["c1"], before["c2"])
    dev.close()


def test_dropin_rotate_example0(gpu):
    """cpsd() then cpsd_rotate() with host arrays, as main() would call them."""
    say("drop-in cpsd_rotate, example0")
    ctl, st = example0_oracle_state()
    cell = cell_from_state(st)
    out = cpsd_amd.cpsd_dropin(cell, 4, ctl["inner"], st.c2.copy(), np.zeros(10), ctl["dt"], ctl["fmass"])
    Er, _, _ = st.step(ctl["inner"], ctl["dt"], ctl["fmass"])
    compare(out["E"], Er)
    rot = cpsd_amd.rotate_dropin(4, cell.n2ft3d, out["hml"], out["c2"])
    ref = O.epilogue(4, cell.n2ft3d, cell.nfft, st.c1, st.c2, st.hml)
    ep = dict(ortho=ref["ortho"], eig=rot["eig"], V=rot["V"])
    check_against_oracle(ep, ref, 4, rot["c1"], out["c2"])


def test_dropin_rotate_needs_a_context(gpu):
    say("drop-in cpsd_rotate size mismatch")
    ctl, st = example0_oracle_state()
    cell = cell_from_state(st)
    cpsd_amd.cpsd_dropin(cell, 4, 1, st.c2.copy(), np.zeros(10), ctl["dt"], ctl["fmass"])
    L = cpsd_amd.hip_lib()
    nea = np.array([5, 0], dtype=np.int32)
    h = np.eye(5).reshape(-1).copy()
    L.cpsd_rotate(1, nea, cell.n2ft3d, h, np.zeros(5), np.zeros(5 * cell.n2ft3d), np.zeros(5 * cell.n2ft3d))
    assert "cpsd_rotate" in cpsd_amd._err()
