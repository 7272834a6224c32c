"""Safeguards of the resident API through the C-ABI: the FP64 refresh period
of the mixed modes, the fallback to the reference's FP64 step when the
energy rises, the non-finite guard and the clean restart after it, and the
precision/ion-motion combinations that are refused.  Every condition is
forced from the host (an unstable preconditioner, an invalid mode); every
kernel still runs to its normal end."""
import sys

import numpy as np
import pytest

import cpsd_amd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_cell():
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    return cell, cell.random_orbitals(8, seed=3, decay=0.05)


def test_refresh_every_step_is_fp64(gpu):
    """mode 1 with an FP64 refresh every step runs only FP64 steps: the bits
    of mode 0; a refresh period of 3 differs from both"""
    say("refresh period")
    cell, psi = small_cell()
    out = {}
    for key, mode, period in (("fp64", 0, 0), ("r1", 1, 1), ("r3", 1, 3), ("r0", 1, 0)):
        dev = cpsd_amd.DeviceSD(cell, 8)
        dev.set_orbitals(psi)
        dev.set_precision(mode)
        dev.set_refresh(period)
        Es = [dev.run(5)[0] for _ in range(3)]
        out[key] = (np.array(Es), dev.get(c2=True)["c2"])
        dev.close()
    assert np.array_equal(out["r1"][0], out["fp64"][0]) and np.array_equal(out["r1"][1], out["fp64"][1])
    assert not np.array_equal(out["r3"][1], out["r0"][1])
    assert not np.array_equal(out["r3"][1], out["fp64"][1])
    assert np.max(np.abs(out["r3"][0] - out["fp64"][0])) < 1e-6


def test_fallback_when_the_energy_rises(gpu):
    """a rise threshold of -10 |E| (every call counts as a rise, a
    condition forced from the host) turns the mixed mode and the
    preconditioner off after the first compared call; the following calls
    are the reference's FP64 step, bit for bit those of a context that was
    switched to mode 0 by hand at the same point"""
    say("fallback")
    cell, psi = small_cell()
    K = np.full(cell.nfft3d, 2.0)
    runs = []
    for forced in (True, False):
        dev = cpsd_amd.DeviceSD(cell, 8)
        dev.set_orbitals(psi)
        dev.set_precision(1)
        dev.set_preconditioner(K)
        if forced:
            dev.set_fallback(True, rise=-10.0)
        Es = [dev.run(5)[0] for _ in range(2)]
        if forced:
            assert dev.fallback_info() == dict(fallbacks=1, mode=0, precond=0)
        else:
            assert dev.fallback_info() == dict(fallbacks=0, mode=1, precond=1)
            dev.set_precision(0)
            dev.set_preconditioner(None)
        Es += [dev.run(5)[0] for _ in range(3)]
        runs.append((np.array(Es), dev.get(c2=True)["c2"]))
        dev.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.all(np.diff(runs[0][0][:, 0]) < 0)


def test_nonfinite_energy_is_an_error_and_new_orbitals_start_clean(gpu):
    """factors large enough to overflow make the call fail with a message;
    new orbitals then run exactly as on a fresh context"""
    say("non-finite guard")
    cell, psi = small_cell()
    dev = cpsd_amd.DeviceSD(cell, 8)
    dev.set_fallback(False)
    dev.set_orbitals(psi)
    dev.run(3)
    dev.set_preconditioner(np.full(cell.nfft3d, 9.0e5))
    with pytest.raises(cpsd_amd.CpsdError, match="non-finite"):
        for _ in range(6):
            dev.run(5)
    dev.set_preconditioner(None)
    dev.set_orbitals(psi)
    a = [dev.run(5)[0] for _ in range(2)]
    ca = dev.get(c2=True)["c2"]
    dev.close()
    fresh = cpsd_amd.DeviceSD(cell, 8)
    fresh.set_orbitals(psi)
    b = [fresh.run(5)[0] for _ in range(2)]
    cb = fresh.get(c2=True)["c2"]
    fresh.close()
    assert np.array_equal(np.array(a), np.array(b)) and np.array_equal(ca, cb)


def test_bf16_mode_refused_with_moving_ions(gpu):
    say("mode 2 with moving ions")
    cell, psi = small_cell()
    dev = cpsd_amd.DeviceSD(cell, 8)
    dev.set_orbitals(psi)
    dev.set_ions(1)
    with pytest.raises(cpsd_amd.CpsdError, match="mode 2"):
        dev.set_precision(2)
    dev.set_ions(0)
    dev.set_precision(2)
    with pytest.raises(cpsd_amd.CpsdError, match="mode 0 or 1"):
        dev.set_ions(1)
    dev.close()
