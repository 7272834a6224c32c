"""GPU parity at the shapes of the benchmark workloads: the grids of the
32-water (96^3) and 64-water (128^3) cells through the whole SD step against
the oracle, orbital counts above 128 (the 256-orbital tile grid of the
symmetric Gram with its diagonal pairs, an odd tile count, the general
Lagrange kernel, four column tiles of the tall kernels) at small grids where
the oracle finishes in seconds, and the full 96^3 x 128-orbital workload of
bench.py through size-independent properties."""
import sys

import numpy as np
import pytest

import cpsd_amd
import oracle_lib as O
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-shapes] " + msg, flush=True)
    sys.stdout.flush()


def run_pair(nmol, L, nfft, ne, steps, inner, seed=7, decay=0.05):
    cell = cpsd_amd.water_cell(nmol, L, nfft, seed=seed)
    assert int(cell.masker[0::2].sum()) >= ne
    psi = cell.random_orbitals(ne, seed=seed + 1, decay=decay)
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl,
                 cell.vnlnrm, cell.r, psi, cell.amass)
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_orbitals(psi)
    for _ in range(steps):
        Eg, deg, dcg = dev.run(inner)
        Er, der, dcr = st.step(inner)
        compare(Eg, Er, tol=1e-9)
        assert abs(deg - der) < 1e-9
    o = dev.get(c2=True, dn=True, hml=True, lmd=True)
    assert np.max(np.abs(o["c2"] - st.c2)) < 1e-10
    assert np.max(np.abs(o["dn"] - st.dn)) < 1e-10
    assert np.max(np.abs(o["hml"] - st.hml)) < 1e-9
    assert np.max(np.abs(o["lmd"] - st.lmd)) < 1e-9
    dev.close()


def test_ne256_vs_oracle(gpu):
    """256 orbitals: 4 x 4 Gram tiles (6 above the diagonal, 2 diagonal
    pairs), general Lagrange kernel, 4 column tiles in the tall kernels"""
    say("ne=256 at 24^3")
    run_pair(2, 12.0, 24, 256, steps=1, inner=1)


@pytest.mark.parametrize("ne", [384, 512])
def test_ne384_512_vs_oracle(gpu, ne):
    """orbital counts above 256: the 32-wide Lagrange kernel (its grid of
    12 x 12 and 16 x 16 workgroups checked for residency), the generic
    carried-overlap kernel (the carry runs from the call's second step), and
    a 2 ne = 768 / 1024 Gram with its split-K partial buffers; a 20^3 grid
    keeps the oracle's O(ne^2) overlaps under a minute"""
    say("ne=%d at 20^3" % ne)
    run_pair(2, 12.0, 20, ne, steps=1, inner=2)


def test_ne160_vs_oracle(gpu):
    """160 orbitals: 3 Gram tile rows, the second diagonal pair half empty"""
    say("ne=160 at 32^3")
    run_pair(4, 14.0, 32, 160, steps=1, inner=1)


@pytest.mark.parametrize("ne,steps", [(128, 2), (256, 1)])
def test_carried_overlap_vs_oracle(gpu, ne, steps):
    """ne a multiple of 128 over several steps per call: from the second step
    on, <c1|c1> comes from the previous step's overlaps and Lagrange
    multipliers (<c2f|c2f> of the constrained orbitals) and the Gram skips its
    C1 x C1 block; calls continue the carried value"""
    say("carried <c1|c1>, ne=%d at 24^3" % ne)
    run_pair(2, 12.0, 24, ne, steps=steps, inner=3, seed=7)


def test_carried_overlap_reset_by_new_orbitals(gpu):
    """new orbitals drop the carried <c1|c1>: a context that ran, then takes
    new orbitals, gives the bits of a fresh context; the epilogue in between
    runs does not disturb the next run (in both precision modes)"""
    say("carried <c1|c1> reset, ne=128 at 24^3")
    ne = 128
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi_a = cell.random_orbitals(ne, seed=1, decay=0.05)
    psi_b = cell.random_orbitals(ne, seed=2, decay=0.05)
    for mode in (0, 1):
        dev = cpsd_amd.DeviceSD(cell, ne)
        dev.set_precision(mode)
        dev.set_orbitals(psi_a)
        dev.run(4)
        dev.set_orbitals(psi_b)
        E1, _, _ = dev.run(4)
        E2, _, _ = dev.run(4)
        c_used = dev.get(c2=True)["c2"]
        fresh = cpsd_amd.DeviceSD(cell, ne)
        fresh.set_precision(mode)
        fresh.set_orbitals(psi_b)
        F1, _, _ = fresh.run(4)
        F2, _, _ = fresh.run(4)
        assert np.array_equal(E1, F1) and np.array_equal(E2, F2)
        assert np.array_equal(c_used, fresh.get(c2=True)["c2"])
        # an epilogue between runs leaves the following run as it would be
        # without it (the next step's c1 is the unrotated current set); its
        # first step forms the whole Gram again, so agreement is to rounding
        fresh.epilogue()
        G3, _, _ = fresh.run(4)
        H3, _, _ = dev.run(4)
        assert np.max(np.abs(G3 - H3)) < 1e-10
        dev.close()
        fresh.close()


def test_grid96_step_vs_oracle(gpu):
    """the 32-water cell of the h2o-32 workload (96^3, 192 projectors) with
    8 orbitals"""
    say("h2o-32 cell, 96^3, ne=8")
    run_pair(32, 18.6, 96, 8, steps=2, inner=1, seed=0, decay=0.0)


def test_grid128_step_vs_oracle(gpu):
    """the 128^3 grid of the h2o-64 workload (two-half z passes, 133 KB
    planes) with one water and 4 orbitals"""
    say("128^3, ne=4")
    run_pair(1, 20.0, 128, 4, steps=2, inner=1, seed=3)


def test_grid128_vpsi_prefetch_vs_oracle(gpu):
    """128^3 with 16 orbitals: the FP64 V psi pass takes two orbitals per
    workgroup and prefetches the second one's plane into registers while the
    first is transformed (with 4 orbitals above it takes one)"""
    say("128^3, ne=16")
    run_pair(4, 20.0, 128, 16, steps=1, inner=1, seed=3)


def test_full_h2o32_properties(gpu):
    """bench.py's workload (32 waters, 96^3, 128 orbitals) at full size:
    orthonormality (the GPU Gram against a host Gram), hml symmetric, the
    electron count, zero pads and masked-out coefficients, a decreasing
    energy, a converging Lagrange iteration, and bit-identical energies when
    the same steps run twice"""
    say("h2o-32 full size properties")
    ne = 128
    cell = cpsd_amd.water_cell(32, 18.6, 96, seed=0)
    psi = cell.random_orbitals(ne, seed=1)
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_orbitals(psi)
    E1, _, _ = dev.run(10)
    E2, _, _ = dev.run(10)
    assert E2[0] < E1[0]
    it, _ = dev.lagrange_info()
    assert 1 <= it < 30
    o = dev.get(c1=True, c2=True, dn=True, hml=True)
    n, n2 = 96, cell.n2ft3d
    H = o["hml"].reshape(ne, ne)
    assert np.array_equal(H, H.T)
    dv = cell.omega / n ** 3
    assert 2 * o["dn"].sum() * dv == pytest.approx(2 * ne, rel=1e-10)
    assert np.all(o["dn"].reshape(n, n, n + 2)[:, :, n:] == 0)
    mask = cell.masker.astype(bool)
    c2 = o["c2"].reshape(ne, n2)
    assert np.all(c2[:, ~mask] == 0)
    # host Gram of c2 under gcdotc's weights on the masked coefficients
    w = np.full(n2, 2.0)
    w.reshape(n * n, n + 2)[:, 0:2] = 1.0
    cm = c2[:, mask] * np.sqrt(w[mask])
    S = cm @ cm.T
    assert np.max(np.abs(S - np.eye(ne))) < 1e-11
    # the GPU Gram of the step's starting orbitals against the host's
    ep = dev.epilogue()
    c1 = o["c1"].reshape(ne, n2)[:, mask] * np.sqrt(w[mask])
    np.testing.assert_allclose(ep["ortho"], c1 @ c1.T, rtol=0, atol=1e-12)
    dev.close()
    # determinism: the same 20 steps from the same start give the same bits
    dev = cpsd_amd.DeviceSD(cell, ne)
    dev.set_orbitals(psi)
    F1, _, _ = dev.run(10)
    F2, _, _ = dev.run(10)
    assert np.array_equal(F1, E1) and np.array_equal(F2, E2)
    dev.close()


@pytest.mark.parametrize("nfft,L,nmol,ne", [(16, 10.0, 1, 1), (16, 10.0, 1, 5), (24, 12.0, 2, 17),
                                            (24, 12.0, 2, 33), (24, 12.0, 2, 100)])
def test_odd_orbital_counts_vs_oracle(gpu, nfft, L, nmol, ne):
    """one orbital, odd counts and counts just past a 16-row block or a
    32-wide split step, and 100 orbitals (a 7 x 7 grid of Lagrange tiles whose
    MFMA K steps run over zero padding from 100 to 112), against the oracle"""
    say("ne=%d at %d^3" % (ne, nfft))
    run_pair(nmol, L, nfft, ne, steps=2, inner=3)
