"""Split-BF16 overlaps of the FP32 steps (cpsd_dev_set_overlaps) on the GPU.

With the switch on, an FP32 step of modes 1 and 2 forms <c2|c1> and <c2|c2>
(ccpsdv3/cpsd.c:415-437) from the FP64 <c1|c1> carried from the previous step
and split-BF16 MFMA products of its step d = c2 - c1; the Lagrange iteration
and the constraint stay FP64, and the next FP64 step forms the full FP64 Gram.
The reference has no such form, so it is measured against the same mode with
FP64 overlaps (the only difference), against the FP64 oracle and by what
the last FP64 step fixes:
- the orbitals a call returns orthonormal to 1e-12 (the last step's full
  FP64 Gram, as in mode 0);
- those orbitals within 5e-9 of the same mode's with FP64 overlaps, and within
  1e-7 of the oracle's (5e-7 for both with the preconditioner, whose larger
  steps carry a larger d);
- the energies within 5e-7 of |E| of both.  A call's energies are those of
  the orbitals its last step starts from, which the FP32 steps before it
  constrained with the split-BF16 overlaps: from the Gaussian start here,
  where |d| is large (E falls from 1.7e3 Ha by hundreds of Ha per call),
  that is about 1e-7 of |E|; the error shrinks with d, and the converged
  energy of the 64-water cell moves by 1e-12 Ha per atom;
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
    print("[gpu-overlaps] " + msg, flush=True)
    sys.stdout.flush()


def ortho_dev(c, ne, n):
    """max |<c_i|c_j> - delta_ij| of orbitals in the half-complex layout"""
    n2 = c.size // ne
    w = np.full(n2, 2.0)
    w.reshape(n * n, n + 2)[:, 0:2] = 1.0
    c = c.reshape(ne, n2)
    return float(np.max(np.abs((c * w) @ c.T - np.eye(ne))))


def make(cell, ne, psi, ovl, K=None, graphs=False):
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_graphs(graphs)
    dev.set_orbitals(psi)
    dev.set_precision(1)
    dev.set_overlaps(ovl)
    if K is not None:
        dev.set_preconditioner(K)
    return dev


def test_overlaps_setter_checked(gpu):
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    dev = cpsd_amd.DeviceSD(cell, 12)
    with pytest.raises(cpsd_amd.CpsdError):
        dev.set_overlaps(True)
    dev.set_overlaps(False)
    dev.close()


@pytest.mark.parametrize("ne,precond", [(128, False), (256, False), (128, True)])
def test_split_overlaps_vs_fp64_overlaps_and_oracle(gpu, ne, precond):
    """3 calls of 5 steps from a Gaussian start, 2 waters at 24^3: the FP32
    steps after the first of each call run on the split-BF16 overlaps"""
    say("split-BF16 overlaps ne=%d precond=%s" % (ne, precond))
    nfft, L = 24, 12.0
    cell = cpsd_amd.water_cell(2, L, nfft, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    K = tpa_factor(cell.tg, 4.0, 1.0) if precond else None
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    a = make(cell, ne, psi, False, K)
    b = make(cell, ne, psi, True, K)
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
        say("  call %d: E %.10f  |E - E(FP64 overlaps)|/|E| %.2e  |E - E(oracle)|/|E| %.2e"
            % (call, Eb[0], dab, dor))
        assert np.isfinite(deb) and np.isfinite(dcb)
    oa = a.get(c2=True)
    ob = b.get(c2=True, dn=True)
    dc_ab = float(np.max(np.abs(ob["c2"] - oa["c2"])))
    dc_or = float(np.max(np.abs(ob["c2"] - st.c2)))
    dn_or = float(np.max(np.abs(ob["dn"] - st.dn)))
    orth = ortho_dev(ob["c2"], ne, nfft)
    say("  c2 vs FP64 overlaps %.2e, vs oracle %.2e; dn vs oracle %.2e; orthonormality %.2e"
        % (dc_ab, dc_or, dn_or, orth))
    a.close()
    b.close()
    assert orth < 1e-12
    assert worst_ab < 5e-7 and worst_or < 5e-7
    assert dc_ab < (5e-7 if precond else 5e-9)
    assert dc_or < (5e-7 if precond else 1e-7)
    assert dn_or < 1e-7


def test_split_overlaps_graphs_bitwise(gpu):
    """a captured call replays the split-BF16 steps bit for bit"""
    say("split-BF16 overlaps, graphs vs eager")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(128, seed=3, decay=0.05)
    out = []
    for graphs in (False, True):
        dev = make(cell, 128, psi, True, graphs=graphs)
        Es = [dev.run(5)[0].copy() for _ in range(4)]
        out.append((np.array(Es), dev.get(c2=True)["c2"]))
        dev.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
