"""The KB transform on the device (cpsd_kbpp_device, ckbppv3/dkbppv3.c:335-444)
against the host transform and the oracle's restatement, which
test_oracle_physics.py and the Example0 golden energy pin."""
import sys
import time

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-kbpp] " + msg, flush=True)
    sys.stdout.flush()


def transform(fn, p, lmax, nfft3d, G):
    vl, vnl, nrm = np.zeros(nfft3d), np.zeros(9 * nfft3d), np.zeros(9)
    rc = fn(p["zv"], lmax, p["nrho"], p["drho"], p["rho"], p["vp"].copy(), p["wp"], nfft3d, G, vl, vnl, nrm)
    assert rc == 0
    return vl, vnl, nrm


@pytest.mark.parametrize("name,lmax,nfft,L", [("O00.dat", 2, 32, 20.0), ("H00.dat", 1, 32, 20.0),
                                              ("C01.dat", 2, 24, 12.0), ("Fe00.dat", 3, 16, 10.0),
                                              ("O00.dat", 0, 16, 10.0)])
def test_device_kb_matches_host(gpu, name, lmax, nfft, L):
    say("kbpp %s lmax=%d %d^3" % (name, lmax, nfft))
    p = cpsd_amd.read_psp(cpsd_amd.os.path.join(cpsd_amd.DATA, "psp", name))
    lmax = min(lmax, p["lmax"])
    cell = cpsd_amd.Cell(1, L, nfft, [(cpsd_amd.os.path.join(cpsd_amd.DATA, "psp", name), lmax)], [0],
                         [0.0, 0.0, 0.0])
    n = cell.nfft3d
    vp0 = p["vp"].copy()
    dl, dn, dr = transform(cpsd_amd.hip_lib().cpsd_kbpp_device, p, lmax, n, cell.G)
    np.testing.assert_array_equal(p["vp"], vp0)
    hl, hn, hr = transform(cpsd_amd.host_lib().cpsdh_kbpp, p, lmax, n, cell.G)
    sc_l = np.abs(hl).max()
    np.testing.assert_allclose(dl, hl, rtol=0, atol=1e-12 * sc_l)
    m = max(1, lmax * lmax) * n
    sc_n = max(1e-300, np.abs(hn[:m]).max())
    np.testing.assert_allclose(dn[:m], hn[:m], rtol=0, atol=1e-12 * sc_n)
    np.testing.assert_allclose(dr, hr, rtol=1e-13, atol=0)
    # the oracle's restatement of the reference transform
    s = O.Setup(1, L, nfft, np.array([0], dtype=np.int32), np.array([p["zv"]]))
    ol, on, orr = O.kbpp(O.read_psp(cpsd_amd.os.path.join(cpsd_amd.DATA, "psp", name)), lmax, s)
    np.testing.assert_allclose(dl, ol, rtol=0, atol=1e-12 * sc_l)
    np.testing.assert_allclose(dn[:m], on[:m], rtol=0, atol=1e-12 * sc_n)


def test_device_kb_closed_forms(gpu):
    """every channel (lmax 3) of the device transform against the closed-form
    Hankel transforms of Gaussian radial functions (the host test's table)"""
    from test_host_inputs import gaussian_table, kb_closed_forms
    say("kbpp closed forms")
    r, vp, wp, drho, b, a, beta = gaussian_table()
    s = O.Setup(1, 20.0, 16, np.array([0], dtype=np.int32), np.array([1.0]))
    vl, vnl, nrm = np.zeros(s.nfft3d), np.zeros(9 * s.nfft3d), np.zeros(9)
    assert cpsd_amd.hip_lib().cpsd_kbpp_device(0.0, 3, r.size, drho, r, vp, wp, s.nfft3d, s.G, vl, vnl, nrm) == 0
    G = s.G.reshape(3, -1)[:, 1:]
    q = np.sqrt((G * G).sum(0))
    gx, gy, gz = G / q
    cs, cp, cd, cl = kb_closed_forms(q, a, b, 0.0, r[-1])
    V = vnl.reshape(9, s.nfft3d)[:, 1:]
    assert np.max(np.abs(vl[1:] - cl)) < 1e-9 * np.max(np.abs(cl))
    assert np.max(np.abs(V[0] - cs)) < 1e-9 * np.max(np.abs(cs))
    for row, ang in ((1, gx), (2, gy), (3, gz)):
        assert np.max(np.abs(V[row] - cp * ang)) < 1e-9 * np.max(np.abs(cp))
    for row, ang in ((4, (3 * gz * gz - 1) / (2 * np.sqrt(3))), (5, gx * gy), (6, gy * gz), (7, gz * gx),
                     (8, (gx * gx - gy * gy) / 2)):
        assert np.max(np.abs(V[row] - cd * ang)) < 1e-9 * np.max(np.abs(cd))


def test_device_kb_water_cell_runs_the_same(gpu):
    """a 2-water cell built with the device transform gives the same SD energies"""
    say("water cell with device KB")
    a = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    b = cpsd_amd.water_cell(2, 12.0, 24, seed=7, kb_device=True)
    np.testing.assert_allclose(b.vl, a.vl, rtol=0, atol=1e-12 * np.abs(a.vl).max())
    np.testing.assert_allclose(b.vnl, a.vnl, rtol=0, atol=1e-12 * np.abs(a.vnl).max())
    psi = a.random_orbitals(8, seed=3)
    E = []
    for cell in (a, b):
        dev = cpsd_amd.DeviceSD(cell, 8)
        dev.set_orbitals(psi)
        E.append(dev.run(5)[0])
        dev.close()
    np.testing.assert_allclose(E[1], E[0], rtol=0, atol=1e-10)


def test_device_kb_timing_96(gpu):
    """the 96^3 grid of the h2o-32 workload: device and host transforms agree;
    both times are printed"""
    say("kbpp timing 96^3")
    p = cpsd_amd.read_psp(cpsd_amd.os.path.join(cpsd_amd.DATA, "psp", "O00.dat"))
    cell = cpsd_amd.Cell(1, 18.6, 96, [(cpsd_amd.os.path.join(cpsd_amd.DATA, "psp", "O00.dat"), 2)], [0],
                         [0.0, 0.0, 0.0])
    n = cell.nfft3d
    transform(cpsd_amd.hip_lib().cpsd_kbpp_device, p, 2, n, cell.G)      # warm-up
    t0 = time.perf_counter()
    dl, dn, _ = transform(cpsd_amd.hip_lib().cpsd_kbpp_device, p, 2, n, cell.G)
    td = time.perf_counter() - t0
    t0 = time.perf_counter()
    hl, hn, _ = transform(cpsd_amd.host_lib().cpsdh_kbpp, p, 2, n, cell.G)
    th = time.perf_counter() - t0
    say("96^3 O s+p: device %.3f s (incl. copies), host %.3f s" % (td, th))
    np.testing.assert_allclose(dl, hl, rtol=0, atol=1e-12 * np.abs(hl).max())
    np.testing.assert_allclose(dn[:4 * n], hn[:4 * n], rtol=0, atol=1e-12 * np.abs(hn[:4 * n]).max())
