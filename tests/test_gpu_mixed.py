"""The FP32-transform precision mode (cpsd_dev_set_precision(h, 1)) on the GPU.

The reference's FP64 build has no such mode, so it is measured by its error
against FP64: against the FP64 oracle over a fixed number of steps, against
the reference's own Example0 run (E after 1000 steps), and against the FP64
device path on the 32-water workload.

Bars:
- converged energies within 1e-6 Ha per atom of FP64;
- the orbitals within 1e-7 of FP64 after a few steps. One FP32 step moves them
  by dte times the rounding of -H psi, about 1e-8 here;
- orthonormality and the electron count at the FP64 level, since the constraint
  and the density sum stay FP64;
- a one-step call (its only step is the last, FP64 one) bit-identical to FP64;
- with moving ions, positions and forces within bounds set from the measured
  drift of the FP32 steps' forces (mode 1 and mode 2).
"""
import sys

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from example_inputs import example0_expected, example0_oracle_state, water_oracle_state
from test_gpu_parity import cell_from_state

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-mixed] " + msg, flush=True)
    sys.stdout.flush()


def gram(c, ne, n):
    n2 = c.size // ne
    w = np.full(n2, 2.0)
    w.reshape(n * n, n + 2)[:, 0:2] = 1.0
    c = c.reshape(ne, n2)
    return (c * w) @ c.T


def test_mixed_mode_argument_checked(gpu):
    st = water_oracle_state()
    dev = cpsd_amd.DeviceSD(cell_from_state(st), 4)
    with pytest.raises(cpsd_amd.CpsdError):
        dev.set_precision(3)
    dev.close()


def test_mixed_single_step_call_is_fp64(gpu):
    """a call's last step runs in FP64, so inner = 1 gives FP64's bits"""
    say("mixed inner=1 bit-identical to FP64")
    st = water_oracle_state()
    cell = cell_from_state(st)
    out = []
    for mode in (0, 1):
        dev = cpsd_amd.DeviceSD(cell, 4)
        dev.set_precision(mode)
        dev.set_orbitals(st.c2)
        E = [dev.run(1)[0] for _ in range(3)]
        out.append((np.array(E), dev.get(c2=True)["c2"]))
        dev.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_mixed_water_vs_oracle(gpu):
    """one water at 32^3: 4 calls of 5 steps (4 FP32 + 1 FP64 each) against
    the FP64 oracle"""
    say("mixed water 32^3 vs oracle")
    st = water_oracle_state()
    cell = cell_from_state(st)
    dev = cpsd_amd.DeviceSD(cell, 4)
    dev.set_precision(1)
    dev.set_orbitals(st.c2)
    ref64 = cpsd_amd.DeviceSD(cell, 4)
    ref64.set_orbitals(st.c2)
    worst = 0.0
    for _ in range(4):
        Eg, deg, dcg = dev.run(5)
        E64, _, _ = ref64.run(5)
        Er, der, dcr = st.step(5)
        d = np.max(np.abs(Eg - Er))
        worst = max(worst, d)
        assert d < 3e-6, (Eg, Er)          # 1e-6 Ha per atom, 3 atoms
        assert abs(deg - der) < 1e-6
    say("max |E_mixed - E_oracle| over the calls: %.3e" % worst)
    o = dev.get(c2=True, dn=True)
    assert np.max(np.abs(o["c2"] - st.c2)) < 1e-7
    assert np.max(np.abs(o["dn"] - st.dn)) < 1e-7
    # the FP32 transforms did run: the result is not FP64's bit pattern
    assert not np.array_equal(o["c2"], ref64.get(c2=True)["c2"])
    S = gram(o["c2"], 4, cell.nfft)
    assert np.max(np.abs(S - np.eye(4))) < 1e-11
    dev.close()
    ref64.close()


def test_mixed_example0_golden(gpu):
    """the reference's Example0 run (C2, 1000 steps) in the mixed mode against
    its FP64 energy"""
    say("mixed example0 golden")
    exp = example0_expected()
    ctl, st = example0_oracle_state()
    dev = cpsd_amd.DeviceSD(cell_from_state(st), 4)
    dev.set_precision(1)
    dev.set_orbitals(st.c2)
    for _ in range(ctl["outer"]):
        E, de, dc = dev.run(ctl["inner"], ctl["dt"], ctl["fmass"])
    say("E0 mixed %.12f golden %.12f" % (E[0], exp["E0"]))
    assert abs(E[0] - exp["E0"]) < 2e-6          # 1e-6 Ha per atom, 2 atoms
    dev.close()


def test_mixed_h2o32_vs_fp64(gpu):
    """the 32-water workload (96^3, 128 orbitals): mixed against the FP64
    device path from the same start, per call and per atom"""
    say("mixed h2o-32 vs FP64")
    ne, natom = 128, 96
    cell = cpsd_amd.water_cell(32, 18.6, 96, seed=0)
    psi = cell.random_orbitals(ne, seed=1)
    devs = []
    for mode in (0, 1):
        d = cpsd_amd.DeviceSD(cell, ne)
        d.set_precision(mode)
        d.set_orbitals(psi)
        devs.append(d)
    worst = 0.0
    for _ in range(3):
        E64, _, _ = devs[0].run(10)
        E32, _, _ = devs[1].run(10)
        worst = max(worst, abs(E32[0] - E64[0]) / natom)
        assert abs(E32[0] - E64[0]) < 1e-6 * natom, (E32[0], E64[0])
    say("max |E_mixed - E_FP64| per atom: %.3e" % worst)
    o = devs[1].get(c2=True, dn=True)
    S = gram(o["c2"], ne, 96)
    assert np.max(np.abs(S - np.eye(ne))) < 1e-11
    dv = cell.omega / 96 ** 3
    assert 2 * o["dn"].sum() * dv == pytest.approx(2 * ne, rel=1e-10)
    for d in devs:
        d.close()


@pytest.mark.parametrize("nmol,L,nfft,ne", [(1, 20.0, 128, 4), (4, 14.0, 32, 160), (11, 16.0, 32, 24),
                                            (20, 16.0, 32, 24)])
def test_mixed_shapes_vs_fp64(gpu, nmol, L, nfft, ne):
    """the 128^3 transforms (two-half z passes); 160 orbitals (two column
    tiles in both FP32 KB products); 66 projectors (one partial 96-row
    projection tile); 120 projectors (the 128-row projection tile): mixed
    against FP64 from the same start"""
    say("mixed vs FP64: %d waters, %d^3, ne=%d" % (nmol, nfft, ne))
    cell = cpsd_amd.water_cell(nmol, L, nfft, seed=4)
    psi = cell.random_orbitals(ne, seed=9, decay=0.05)
    natom = 3 * nmol
    devs = []
    for mode in (0, 1):
        d = cpsd_amd.DeviceSD(cell, ne)
        d.set_precision(mode)
        d.set_orbitals(psi)
        devs.append(d)
    for _ in range(2):
        E64, _, _ = devs[0].run(5)
        E32, _, _ = devs[1].run(5)
        assert np.max(np.abs(E32 - E64)) < 1e-6 * natom, (E32, E64)
    o64 = devs[0].get(c2=True)["c2"]
    o32 = devs[1].get(c2=True, dn=True)
    assert np.max(np.abs(o32["c2"] - o64)) < 1e-6
    assert not np.array_equal(o32["c2"], o64)
    S = gram(o32["c2"], ne, nfft)
    assert np.max(np.abs(S - np.eye(ne))) < 1e-11
    dv = cell.omega / nfft ** 3
    assert 2 * o32["dn"].sum() * dv == pytest.approx(2 * ne, rel=1e-10)
    for d in devs:
        d.close()


def test_bf16_projection_converges_to_fp64(gpu):
    """mode 2 (BF16-input KB projection, 24 projectors so the MFMA kernel
    runs): converged energy within 1e-6 Ha per atom of FP64's, and the BF16
    kernel did run (mode 2's result differs from mode 1's)"""
    say("mode 2 (BF16 projection) converged vs FP64, 4 waters 32^3 ne=16")
    cell = cpsd_amd.water_cell(4, 14.0, 32, seed=0)
    ne, natom = 16, 12
    psi = cell.random_orbitals(ne, seed=1, decay=0.05)
    out = {}
    for mode in (0, 1, 2):
        d = cpsd_amd.DeviceSD(cell, ne)
        d.set_precision(mode)
        d.set_orbitals(psi)
        Ep = None
        for _ in range(1000):
            E, _, _ = d.run(10)
            if Ep is not None and abs(E[0] - Ep) < 1e-7:
                break
            Ep = E[0]
        out[mode] = E[0]
        o = d.get(c2=True)["c2"]
        S = gram(o, ne, 32)
        assert np.max(np.abs(S - np.eye(ne))) < 1e-11
        d.close()
    say("E fp64 %.12f mode1 %.12f mode2 %.12f" % (out[0], out[1], out[2]))
    assert abs(out[1] - out[0]) < 1e-10 * natom
    assert abs(out[2] - out[0]) < 1e-6 * natom
    assert out[2] != out[1]


@pytest.mark.parametrize("mode,tol_e,tol_r,tol_f", [(1, 1e-6, 1e-6, 1e-6)])
def test_mixed_moving_ions_vs_fp64(gpu, mode, tol_e, tol_r, tol_f):
    """move = 1 in mixed mode 1: 4 waters at 32^3, 16 orbitals and 24
    projectors (so the FP32 projection kernel runs), 4 calls of 5 steps
    against FP64 from the same start.  Every step moves the ions, the FP32
    ones with forces from their FP32 products, so positions and forces drift
    from FP64's.  Measured after 20 steps, with the ions moved about 2 bohr:
    4e-8 bohr and 6e-8 Ha/bohr; the bounds are about ten times those.  A
    missing or wrong force term moves forces by ~0.1 Ha/bohr.  Mode 2 (BF16
    projection coefficients in the nonlocal force) moved forces by 3e-4
    Ha/bohr and is refused with moving ions (test_gpu_safeguards.py)."""
    say("mode %d with moving ions vs FP64, 4 waters 32^3 ne=16" % mode)
    cell = cpsd_amd.water_cell(4, 14.0, 32, seed=0)
    ne, natom = 16, 12
    psi = cell.random_orbitals(ne, seed=1, decay=0.05)
    devs = []
    for m in (0, mode):
        d = cpsd_amd.DeviceSD(cell, ne)
        d.set_precision(m)
        d.set_ions(1)
        d.set_orbitals(psi)
        devs.append(d)
    for _ in range(4):
        E64, _, _ = devs[0].run(5)
        Em, _, _ = devs[1].run(5)
        a, b = devs[0].get_ions(), devs[1].get_ions()
        assert abs(Em[0] - E64[0]) < tol_e * natom, (Em[0], E64[0])
        assert np.max(np.abs(a["r2"] - b["r2"])) < tol_r
        assert np.max(np.abs(a["fion"] - b["fion"])) < tol_f
    assert np.max(np.abs(a["r2"] - cell.r)) > 0.5        # the ions did move
    assert not np.array_equal(a["r2"], b["r2"])         # the FP32 steps did run
    for d in devs:
        d.close()
