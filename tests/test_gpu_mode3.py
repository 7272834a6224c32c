"""Precision mode 3 (cpsd_dev_set_precision(h, 3)) on the GPU: FP32 orbital
FFTs and every product of the FP32 steps in split BF16 MFMA (each FP32 factor
as hi + lo BF16, three products, FP32 sums): the KB projection and apply, the
preconditioned update, the constraint, and the overlaps (ccpsdv3/cpsd.c:
415-437), which are formed from the FP64 <c1|c1> carried from the previous
step and the products of the step d = c2 - c1.  The Lagrange iteration stays
FP64 and the next FP64 step (the last of every call) forms the full FP64
Gram.

The reference has no such mode, so it is measured against mode 1 (FP32
products, FP64 overlaps and constraint) and the FP64 oracle from the same
start, and at convergence against the FP64 minimum:
- the orbitals a call returns orthonormal to 1e-12 (the last step's full FP64
  Gram, as in mode 0);
- those orbitals within 1e-6 of the oracle's and the density within 5e-7
  (measured 1.4-5.2e-7 and 8e-8 from this Gaussian start: the split
  products' 2^-16 per term against FP32's 2^-24, in every step of the call
  but its last);
- the energies within 5e-7 of |E| of both.  A call's energies are those of
  the orbitals its last step starts from, which the FP32 steps before it
  constrained with the split-BF16 overlaps: from the Gaussian start here,
  where |d| is large (E falls from 1.7e3 Ha by hundreds of Ha per call), that
  is about 1e-7 of |E|; the error shrinks with d;
- the same bars with an FP64 refresh step every 3 steps inside the calls
  (measured: energies 1-6e-8 of |E|, c2 1.3e-7, dn 6.5e-8 from the oracle);
- with the learned preconditioner on the 32-water cell, the converged energy
  within 1e-8 Ha per atom of the FP64 run's (same preconditioner, same start);
- graph replays bit-identical to eager calls.
"""
import sys

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from test_gpu_precond import tpa_factor

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-mode3] " + msg, flush=True)
    sys.stdout.flush()


def ortho_dev(c, ne, n):
    """max |<c_i|c_j> - delta_ij| of orbitals in the half-complex layout"""
    n2 = c.size // ne
    w = np.full(n2, 2.0)
    w.reshape(n * n, n + 2)[:, 0:2] = 1.0
    c = c.reshape(ne, n2)
    return float(np.max(np.abs((c * w) @ c.T - np.eye(ne))))


def make(cell, ne, psi, mode, K=None, graphs=False):
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_graphs(graphs)
    dev.set_orbitals(psi)
    dev.set_precision(mode)
    if K is not None:
        dev.set_preconditioner(K)
    return dev


def test_mode3_checked(gpu):
    """mode 3 needs 128 or 256 orbitals and fixed ions"""
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    dev = cpsd_amd.DeviceSD(cell, 12)
    with pytest.raises(cpsd_amd.CpsdError):
        dev.set_precision(3)
    dev.set_precision(1)
    dev.close()
    dev = cpsd_amd.DeviceSD(cell, 128)
    dev.set_precision(3)
    with pytest.raises(cpsd_amd.CpsdError):
        dev.set_ions(1, cell.r, cell.amass)
    dev.set_precision(0)
    dev.close()


@pytest.mark.parametrize("ne,precond", [(128, False), (256, False), (128, True)])
def test_mode3_vs_mode1_and_oracle(gpu, ne, precond):
    """3 calls of 5 steps from a Gaussian start, 2 waters at 24^3"""
    say("mode 3 ne=%d precond=%s" % (ne, precond))
    nfft, L = 24, 12.0
    cell = cpsd_amd.water_cell(2, L, nfft, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    K = tpa_factor(cell.tg, 4.0, 1.0) if precond else None
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    a = make(cell, ne, psi, 1, K)
    b = make(cell, ne, psi, 3, K)
    h = None
    worst_ab = worst_or = 0.0
    for call in range(3):
        Ea, _, _ = a.run(5)
        Eb, deb, dcb = b.run(5)
        Er, _, _ = st.step(5, K=None if (h is None or K is None) else K, h=h)
        h = st.hml.copy()
        scale = max(1.0, abs(Ea[0]))
        dab = float(np.max(np.abs(Eb[:10] - Ea[:10]))) / scale
        dor = float(np.max(np.abs(Eb[:10] - Er[:10]))) / scale
        worst_ab, worst_or = max(worst_ab, dab), max(worst_or, dor)
        say("  call %d: E %.10f  |E - E(mode 1)|/|E| %.2e  |E - E(oracle)|/|E| %.2e" % (call, Eb[0], dab, dor))
        assert np.isfinite(deb) and np.isfinite(dcb)
    oa = a.get(c2=True)
    ob = b.get(c2=True, dn=True)
    dc_ab = float(np.max(np.abs(ob["c2"] - oa["c2"])))
    dc_or = float(np.max(np.abs(ob["c2"] - st.c2)))
    dn_or = float(np.max(np.abs(ob["dn"] - st.dn)))
    orth = ortho_dev(ob["c2"], ne, nfft)
    say("  c2 vs mode 1 %.2e, vs oracle %.2e; dn vs oracle %.2e; orthonormality %.2e"
        % (dc_ab, dc_or, dn_or, orth))
    a.close()
    b.close()
    assert orth < 1e-12
    assert worst_ab < 5e-7 and worst_or < 5e-7
    assert dc_or < 1e-6
    assert dn_or < 5e-7


def test_mode3_refresh_period(gpu):
    """an FP64 step every 3 steps inside mode-3 calls (cpsd_dev_set_refresh):
    the refresh steps form the full FP64 Gram after split-BF16 steps and the
    steps after them carry its <c1|c1>; the same bars against the oracle"""
    say("mode 3 with an FP64 refresh every 3 steps")
    nfft, L, ne = 24, 12.0, 128
    cell = cpsd_amd.water_cell(2, L, nfft, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    dev = make(cell, ne, psi, 3)
    dev.set_refresh(3)
    worst = 0.0
    for call in range(3):
        Eb, deb, dcb = dev.run(5)
        Er, _, _ = st.step(5)
        d = float(np.max(np.abs(Eb[:10] - Er[:10]))) / max(1.0, abs(Er[0]))
        worst = max(worst, d)
        say("  call %d: E %.10f  |E - E(oracle)|/|E| %.2e" % (call, Eb[0], d))
        assert np.isfinite(deb) and np.isfinite(dcb)
    ob = dev.get(c2=True, dn=True)
    dc = float(np.max(np.abs(ob["c2"] - st.c2)))
    dn = float(np.max(np.abs(ob["dn"] - st.dn)))
    orth = ortho_dev(ob["c2"], ne, nfft)
    say("  c2 vs oracle %.2e; dn vs oracle %.2e; orthonormality %.2e" % (dc, dn, orth))
    dev.close()
    assert orth < 1e-12
    assert worst < 5e-7
    assert dc < 1e-6 and dn < 5e-7


def test_mode3_graphs_bitwise(gpu):
    """a captured call replays the split-BF16 steps bit for bit"""
    say("mode 3, graphs vs eager")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(128, seed=3, decay=0.05)
    K = tpa_factor(cell.tg, 4.0, 1.0)
    out = []
    for graphs in (False, True):
        dev = make(cell, 128, psi, 3, K, graphs=graphs)
        Es = [dev.run(5)[0].copy() for _ in range(4)]
        out.append((np.array(Es), dev.get(c2=True)["c2"]))
        dev.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_mode3_learned_converges_to_fp64_minimum_h2o32(gpu):
    """the 32-water cell (96^3, 128 orbitals) from its Gaussian start with the
    committed learned preconditioner, FP64 and mode 3, each to |dE| < 1e-7 Ha
    between evaluations: the same minimum within 1e-8 Ha per atom, in about
    the same number of steps, with no fallback"""
    import learned
    say("mode 3 + learned converged vs FP64 + learned, h2o-32")
    model = learned.load_model()
    cell = cpsd_amd.water_cell(32, 18.6, 96, seed=0)
    psi = cell.random_orbitals(128, seed=1, decay=0.0)
    dev = cpsd_amd.DeviceSD(cell, 128)
    dev.set_learned_preconditioner(model["params"], model["alpha"], model["kmin"], model["kmax"])
    res = {}
    for mode in (0, 3):
        dev.set_precision(mode)
        res[mode] = learned.converge(dev, psi)
        say("  mode %d: converged %s in %d steps, E %.10f" % (mode, res[mode][0], res[mode][1], res[mode][2]))
    fb = dev.fallback_info()["fallbacks"]
    dev.close()
    assert res[0][0] and res[3][0] and fb == 0
    assert abs(res[3][2] - res[0][2]) / 96 < 1e-8
    assert res[3][1] <= 1.2 * res[0][1]
