"""Moving ions (move = 1) on the GPU against the oracle: ionic forces
(ccpsdv3/cpsd.c:246-364), the ion update (:401-411) and the SD run that
follows the moving ions, through the drop-in cpsd() symbol and the resident
cpsd_dev_* API.  The oracle's forces are pinned by test_oracle_forces.py."""
import sys

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from example_inputs import example0_oracle_state, water_oracle_state
from test_gpu_parity import cell_from_state, compare

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-forces] " + msg, flush=True)
    sys.stdout.flush()


def test_dropin_moving_ions_example0(gpu):
    """cpsd(move=1) on the Example0 C2 cell, 3 calls of 10 steps, as
    ccpsdv3/main.c:422-426 calls it with the positions carried over."""
    say("dropin move=1 example0")
    ctl, st = example0_oracle_state()
    cell = cell_from_state(st)
    c2, r2 = st.c2.copy(), st.r2.copy()
    E = np.zeros(10)
    for call in range(3):
        out = cpsd_amd.cpsd_dropin(cell, 4, ctl["inner"], c2, E, ctl["dt"], ctl["fmass"], move=1, r2=r2)
        Er, der, dcr = st.step(ctl["inner"], ctl["dt"], ctl["fmass"], move=1)
        compare(out["E"], Er)
        assert abs(out["deltae"] - der) < 1e-9
        assert np.max(np.abs(out["fion"] - st.fion)) < 1e-9
        assert np.max(np.abs(out["r1"] - st.r1)) < 1e-12
        assert np.max(np.abs(out["r2"] - st.r2)) < 1e-12
        assert out["deltar"] == pytest.approx(st.deltar.value, rel=1e-9, abs=1e-12)
        assert np.max(np.abs(out["c2"] - st.c2)) < 1e-10
        assert np.max(np.abs(out["dn"] - st.dn)) < 1e-10
        c2, r2, E = out["c2"], out["r2"], out["E"]
    assert np.max(np.abs(r2 - cell.r)) > 1e-6   # the ions did move


def test_device_moving_ions_water(gpu):
    """resident API: water molecule (O s+p, H s projectors) at 32^3"""
    say("device move=1 water")
    st = water_oracle_state()
    cell = cell_from_state(st)
    dev = cpsd_amd.DeviceSD(cell, 4)
    dev.set_ions(1)
    dev.set_orbitals(st.c2)
    for _ in range(3):
        Eg, deg, dcg = dev.run(5)
        Er, der, dcr = st.step(5, move=1)
        compare(Eg, Er)
        ions = dev.get_ions()
        assert np.max(np.abs(ions["fion"] - st.fion)) < 1e-9
        assert np.max(np.abs(ions["r1"] - st.r1)) < 1e-12
        assert np.max(np.abs(ions["r2"] - st.r2)) < 1e-12
        assert dev.deltar == pytest.approx(st.deltar.value, rel=1e-9)
    o = dev.get(c2=True, dn=True)
    assert np.max(np.abs(o["c2"] - st.c2)) < 1e-10
    dev.close()


def test_device_moving_ions_multi_water(gpu):
    """2 waters at 24^3 with 12 orbitals from the product's input builder"""
    say("device move=1 two waters 24^3")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(12, seed=3, decay=0.05)
    st = O.State(O.Setup(1, 12.0, 24, cell.katm, cell.zv), 12, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    dev = cpsd_amd.DeviceSD(cell, 12)
    dev.set_ions(1)
    dev.set_orbitals(psi)
    for _ in range(2):
        Eg, _, _ = dev.run(3)
        Er, _, _ = st.step(3, move=1)
        compare(Eg, Er)
        ions = dev.get_ions()
        assert np.max(np.abs(ions["fion"] - st.fion)) < 1e-9
        assert np.max(np.abs(ions["r2"] - st.r2)) < 1e-12
    dev.close()


def test_fixed_ions_leave_positions(gpu):
    """move = 0 after set_ions(0): positions stay, no forces are reported"""
    say("device move=0 positions")
    st = water_oracle_state()
    cell = cell_from_state(st)
    dev = cpsd_amd.DeviceSD(cell, 4)
    dev.set_ions(0)
    dev.set_orbitals(st.c2)
    Eg, _, _ = dev.run(3)
    Er, _, _ = st.step(3)
    compare(Eg, Er)
    ions = dev.get_ions()
    assert np.array_equal(ions["r1"], cell.r) and np.array_equal(ions["r2"], cell.r)
    assert dev.deltar == 0.0
    dev.close()
