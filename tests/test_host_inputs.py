"""The product's host input builder (libcpsd_host.so) against the oracle's
restatement of the reference setup, KB generator and plane-wave start."""
import os

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O


@pytest.mark.parametrize("icube,unit,nfft", [(1, 20.0, 32), (1, 18.6, 24), (2, 14.0, 16), (3, 12.0, 12)])
def test_setup_tables_match_oracle(icube, unit, nfft):
    katm = np.array([0, 1, 1], dtype=np.int32)
    zv = np.array([6.0, 1.0])
    ref = O.Setup(icube, unit, nfft, katm, zv)
    c = cpsd_amd.Cell.__new__(cpsd_amd.Cell)
    # run only the setup part of Cell through the host library
    import ctypes as C
    L = cpsd_amd.host_lib()
    n3 = ref.nfft3d
    om, cew, rc = C.c_double(), C.c_double(), C.c_double()
    ua, ug, G = np.zeros(9), np.zeros(9), np.zeros(3 * n3)
    m = np.zeros(ref.n2ft3d, dtype=np.int32)
    tg, vc, vg, rcell = np.zeros(n3), np.zeros(n3), np.zeros(n3), np.zeros(81)
    assert L.cpsdh_setup(icube, unit, nfft, 3, katm, zv, C.byref(om), ua, ug, G, m, tg, vc, vg,
                         rcell, C.byref(cew), C.byref(rc)) == 0
    assert np.array_equal(m, ref.masker)
    for a, b in ((ua, ref.unita), (ug, ref.unitg), (G, ref.G), (tg, ref.tg), (vc, ref.vc),
                 (vg, ref.vg), (rcell, ref.rcell)):
        assert np.allclose(a, b, rtol=1e-15, atol=1e-15)
    assert om.value == pytest.approx(ref.omega, rel=1e-15)
    assert cew.value == pytest.approx(ref.cewald, rel=1e-13)
    assert rc.value == ref.rcut


def test_psp_reader_accepts_fourth_header_token():
    # O00.dat line 2 is "6. 16.00 2 1": the 4th token must not shift rc/nrho
    p = cpsd_amd.read_psp(os.path.join(cpsd_amd.DATA, "psp", "O00.dat"))
    assert (p["zv"], p["amass"], p["lmax"], p["nrho"]) == (6.0, 16.0, 2, 501)
    assert p["drho"] == pytest.approx(0.02)
    assert np.allclose(p["rc"], [0.515471, 0.319391, 0.515471])
    q = O.read_psp(os.path.join(O.DATA, "psp", "O00.dat"))
    assert np.array_equal(p["vp"], q["vp"]) and np.array_equal(p["wp"], q["wp"])
    c = cpsd_amd.read_psp(os.path.join(cpsd_amd.DATA, "psp", "C01.dat"))
    assert (c["zv"], c["lmax"], c["nrho"]) == (4.0, 2, 501)


@pytest.mark.parametrize("name,lmax", [("C01", 2), ("O00", 2), ("H00", 1), ("O00", 1)])
def test_kbpp_matches_oracle(name, lmax):
    s = O.Setup(1, 20.0, 24, [0], [1.0])
    path = os.path.join(O.DATA, "psp", name + ".dat")
    q = O.read_psp(path)
    vl0, vnl0, n0 = O.kbpp(q, lmax, s)
    p = cpsd_amd.read_psp(path)
    vl, vnl, nrm = np.zeros(s.nfft3d), np.zeros(9 * s.nfft3d), np.zeros(9)
    assert cpsd_amd.host_lib().cpsdh_kbpp(p["zv"], lmax, p["nrho"], p["drho"], p["rho"], p["vp"],
                                          p["wp"], s.nfft3d, s.G, vl, vnl, nrm) == 0
    lm = lmax * lmax
    assert np.allclose(vl, vl0, rtol=1e-14, atol=1e-14)
    assert np.allclose(vnl[: lm * s.nfft3d], vnl0[: lm * s.nfft3d], rtol=1e-14, atol=1e-14)
    assert np.allclose(nrm[:lm], n0[:lm], rtol=1e-14)


def test_makepsi_matches_oracle():
    spec = O.read_psi0(os.path.join(O.DATA, "PSI0.DAT"), 40)
    ref = O.makepsi(16, 40, spec)
    got = np.zeros_like(ref)
    cpsd_amd.host_lib().cpsdh_makepsi(16, 40, spec, got)
    assert np.array_equal(got, ref)


def test_random_orbitals_orthonormal_hermitian_masked():
    cell = cpsd_amd.water_cell(1, 12.0, 16, seed=3)
    ne = 6
    psi = cell.random_orbitals(ne, seed=11)
    n2 = cell.n2ft3d
    S = np.array([[O.gcdotc(16, psi[i * n2:(i + 1) * n2], psi[j * n2:(j + 1) * n2])
                   for j in range(ne)] for i in range(ne)])
    assert np.max(np.abs(S - np.eye(ne))) < 1e-13
    m = np.tile(cell.masker, ne)
    assert np.all(psi[m == 0] == 0)
    # the kx = 0 plane is Hermitian, so the real-space orbital is real:
    # c2r of the orbital equals its Hermitian-consistent transform
    p0 = psi[:n2].reshape(16, 16, 18)
    for ky in range(16):
        for kz in range(16):
            a = p0[kz, ky, 0:2]
            b = p0[(16 - kz) % 16, (16 - ky) % 16, 0:2]
            assert a[0] == b[0] and a[1] == -b[1]
    same = np.array_equal(psi, cell.random_orbitals(ne, seed=11))
    assert same
