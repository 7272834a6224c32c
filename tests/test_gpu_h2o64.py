"""The full 64-water workload (128^3 grid, 256 orbitals, 384 projectors: the
cell the metric's target is quoted on) through the resident C-ABI, checked
by size-independent properties: orthonormality of the orbitals (the GPU
Gram of the run epilogue), the electron count, a symmetric hml, decreasing
energy, bit-identical reruns, and the energy of the mixed modes and of the
learned preconditioner against FP64 after the same steps."""
import os
import sys

import numpy as np
import pytest

import cpsd_amd

pytestmark = pytest.mark.gpu

NMOL, L, NFFT, NE = 64, 23.5, 128, 256


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def h2o64(gpu):
    say("h2o-64 setup")
    cell = cpsd_amd.water_cell(NMOL, L, NFFT, seed=3)
    psi = cell.random_orbitals(NE, seed=4, decay=0.0)
    dev = cpsd_amd.DeviceSD(cell, NE)
    yield cell, psi, dev
    dev.close()


def run_calls(dev, psi, calls, mode=0, learned=None):
    dev.set_precision(mode)
    dev.set_preconditioner(None)
    if learned is not None:
        dev.set_learned_preconditioner(learned["params"], learned["alpha"], learned["kmin"], learned["kmax"])
    dev.set_orbitals(psi)
    dev.E[:] = 0.0
    Es = [dev.run(10)[0] for _ in range(calls)]
    dev.set_precision(0)
    dev.set_preconditioner(None)
    return np.array(Es)


def test_h2o64_fp64_properties(h2o64):
    cell, psi, dev = h2o64
    say("h2o-64 fp64 30 steps")
    sz = dev.sizes()
    assert (sz["ng"], sz["nproj"]) == (555380, 384)
    Es = run_calls(dev, psi, 3)
    assert Es[1][0] < Es[0][0] and Es[2][0] < Es[1][0]
    o = dev.get(dn=True, hml=True)
    H = o["hml"].reshape(NE, NE)
    assert np.array_equal(H, H.T)
    dv = cell.omega / NFFT ** 3
    assert 2 * o["dn"].sum() * dv == pytest.approx(2 * NE, rel=1e-10)
    assert np.all(o["dn"].reshape(NFFT, NFFT, NFFT + 2)[:, :, NFFT:] == 0)
    it, ad = dev.lagrange_info()
    assert 1 <= it < 30
    # the same 30 steps again: the same bits
    Es2 = run_calls(dev, psi, 3)
    assert np.array_equal(Es, Es2)
    ep = dev.epilogue()
    assert np.max(np.abs(ep["ortho"] - np.eye(NE))) < 1e-11


def test_h2o64_mixed_modes_vs_fp64(h2o64):
    cell, psi, dev = h2o64
    say("h2o-64 mixed modes")
    E0 = run_calls(dev, psi, 3)
    for mode, bound in ((1, 1e-6), (2, 1e-5)):
        Em = run_calls(dev, psi, 3, mode=mode)
        assert not np.array_equal(Em, E0)          # the reduced-precision kernels ran
        err = np.max(np.abs(Em[:, 0] - E0[:, 0])) / (3 * NMOL)
        say("mode %d: max |E - E_fp64| per atom over 30 steps %.2e" % (mode, err))
        assert err < bound
        ep = dev.epilogue()
        assert np.max(np.abs(ep["ortho"] - np.eye(NE))) < 1e-11


def test_h2o64_learned_preconditioner_steps(h2o64):
    """the committed learned preconditioner in mode 1: 40 steps lower the
    energy below plain FP64 SD's after the same 40 steps, keep orthonormality
    and never fall back"""
    import learned
    cell, psi, dev = h2o64
    say("h2o-64 learned")
    model = learned.load_model()
    E0 = run_calls(dev, psi, 4)
    dev.set_fallback(True)
    El = run_calls(dev, psi, 4, mode=1, learned=model)
    assert dev.fallback_info()["fallbacks"] == 0
    assert np.all(np.diff(El[:, 0]) < 0)
    assert El[-1][0] < E0[-1][0]
    ep = dev.epilogue()
    assert np.max(np.abs(ep["ortho"] - np.eye(NE))) < 1e-11
