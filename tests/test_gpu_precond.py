"""The preconditioned SD step (cpsd_dev_set_preconditioner) on the GPU against
the oracle's preconditioned step (oracle_set_preconditioner), through the
C-ABI, and what it converges to.

Rule both follow: the first call after new orbitals runs the reference's
steps; every later call runs c2 = c1 + K o dte (cpsi + c1 h) in each step,
with h the hml of the previous call's energy evaluation.  Bars: energies
1e-9 Ha, c2 and dn 1e-10 (FP64 on both sides); the converged energy of the
preconditioned run equals plain SD's to the convergence tolerance."""
import sys

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from example_inputs import water_oracle_state
from test_gpu_parity import cell_from_state, compare

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tpa_factor(tg, kmax, e0):
    """a smooth per-G factor of the kinetic-energy family: kmax at small
    1/2 G^2, falling as kmax e0 / G^2 beyond e0"""
    x = tg / e0
    p = 27.0 + 18.0 * x + 12.0 * x ** 2 + 8.0 * x ** 3
    return kmax * p / (p + 16.0 * x ** 4)


def run_pair(dev, st, K, calls, inner):
    h = None
    for call in range(calls):
        Eg, deg, dcg = dev.run(inner)
        Er, der, dcr = st.step(inner, K=None if h is None else K, h=h)
        compare(Eg, Er)
        assert abs(deg - der) < 1e-9
        h = st.hml.copy()
    o = dev.get(c2=True, dn=True, hml=True, lmd=True)
    assert np.max(np.abs(o["dn"] - st.dn)) < 1e-10
    assert np.max(np.abs(o["c2"] - st.c2)) < 1e-10
    assert np.max(np.abs(o["hml"] - st.hml)) < 1e-9
    assert np.max(np.abs(o["lmd"] - st.lmd)) < 1e-9


def test_precond_water_vs_oracle(gpu):
    say("precond water 32^3")
    st = water_oracle_state()
    cell = cell_from_state(st)
    K = tpa_factor(st.s.tg, 8.0, 1.0)
    dev = cpsd_amd.DeviceSD(cell, 4)
    dev.set_orbitals(st.c2)
    dev.set_preconditioner(K)
    run_pair(dev, st, K, 4, 5)
    dev.close()


@pytest.mark.parametrize("nfft,L,nmol,ne,calls,inner", [(24, 12.0, 2, 12, 3, 4), (24, 12.0, 2, 128, 2, 3)])
def test_precond_random_start_vs_oracle(gpu, nfft, L, nmol, ne, calls, inner):
    """Gaussian start; 128 orbitals runs the tiled Gram with the carried
    <c1|c1> and the tall MFMA kernels in every step"""
    say("precond random start N=%d ne=%d" % (nfft, ne))
    cell = cpsd_amd.water_cell(nmol, L, nfft, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    K = tpa_factor(cell.tg, 4.0, 1.0)
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_orbitals(psi)
    dev.set_preconditioner(K)
    run_pair(dev, st, K, calls, inner)
    dev.close()


def test_precond_graphs_bitwise(gpu):
    """captured calls replay the preconditioned steps bit for bit, and a new
    set of factors in the same buffer is picked up by the replay"""
    say("precond graphs")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(12, seed=3, decay=0.05)
    out = []
    for graphs in (False, True):
        dev = cpsd_amd.DeviceSD(cell, 12)
        dev.set_graphs(graphs)
        dev.set_orbitals(psi)
        dev.set_preconditioner(tpa_factor(cell.tg, 4.0, 1.0))
        Es = [dev.run(5)[0] for _ in range(4)]
        dev.set_preconditioner(tpa_factor(cell.tg, 2.0, 1.0))
        Es += [dev.run(5)[0] for _ in range(3)]
        out.append((np.array(Es), dev.get(c2=True)["c2"]))
        dev.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_precond_converges_to_plain_sd_minimum(gpu):
    """4 waters, 32^3, 16 orbitals: run plain and preconditioned SD until
    |dE| < 1e-10 Ha between evaluations; the energies agree to 1e-7 Ha (the
    remaining error of the slower plain run), the preconditioned run takes
    fewer steps, and its orbitals are orthonormal"""
    say("precond converges")
    cell = cpsd_amd.water_cell(4, 14.0, 32, seed=5)
    psi = cell.random_orbitals(16, seed=9, decay=0.05)
    res = {}
    for name, K in (("plain", None), ("precond", tpa_factor(cell.tg, 8.0, 1.0))):
        dev = cpsd_amd.DeviceSD(cell, 16)
        dev.set_orbitals(psi)
        if K is not None:
            dev.set_preconditioner(K)
        Eprev, n = None, 0
        while n < 40000:
            E, _, _ = dev.run(10)
            n += 10
            if Eprev is not None and abs(E[0] - Eprev) < 1e-10:
                break
            Eprev = E[0]
        c2 = dev.get(c2=True)["c2"].reshape(16, cell.n2ft3d)
        w = np.full(cell.n2ft3d, 2.0)
        w.reshape(32 * 32, 34)[:, 0:2] = 1.0
        S = (c2 * w) @ c2.T
        assert np.max(np.abs(S - np.eye(16))) < 1e-11
        res[name] = (n, E[0])
        dev.close()
    say("plain %d steps E %.10f ; precond %d steps E %.10f" % (res["plain"] + res["precond"]))
    assert res["precond"][0] < res["plain"][0] / 2
    assert abs(res["precond"][1] - res["plain"][1]) < 1e-7


def test_precond_in_mixed_mode_tracks_fp64(gpu):
    """mode 1 forms the preconditioned step's c1 h product in FP32 in its
    FP32 steps (the FFTs and KB products are FP32 there too): the energies
    stay within 1e-7 |E| of the FP64 preconditioned run (this start is far
    from converged, E is 450-1700 Ha; measured 1.4e-8 |E|), and the bits
    differ (the FP32 kernels ran)"""
    say("precond mixed mode")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(128, seed=3, decay=0.05)
    K = tpa_factor(cell.tg, 4.0, 1.0)
    out = []
    for mode in (0, 1):
        dev = cpsd_amd.DeviceSD(cell, 128)
        dev.set_orbitals(psi)
        dev.set_precision(mode)
        dev.set_preconditioner(K)
        out.append(np.array([dev.run(5)[0] for _ in range(4)]))
        dev.close()
    assert not np.array_equal(out[0], out[1])
    assert np.max(np.abs(out[0][:, 0] - out[1][:, 0]) / np.abs(out[0][:, 0])) < 1e-7
