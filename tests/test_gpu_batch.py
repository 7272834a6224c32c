"""Several independent cells advanced together (cpsd_dev_run_start / _finish,
cpsd_amd.run_batch) against the same cells run one after the other, bit for
bit: each cell has its own stream and buffers, so running them concurrently
must not change any result (with and without HIP-graph replay, in FP64 and in
the mixed precision modes)."""
import sys

import numpy as np
import pytest

import cpsd_amd

pytestmark = pytest.mark.gpu

CELLS = [(1, 12.0, 24), (2, 13.4, 24), (3, 14.9, 32), (5, 17.7, 48)]


def say(msg):
    print("[gpu-batch] " + msg, flush=True)
    sys.stdout.flush()


def make(graphs, mode=0):
    devs = []
    for i, (nmol, L, n) in enumerate(CELLS):
        cell = cpsd_amd.water_cell(nmol, L, n, seed=i)
        d = cpsd_amd.DeviceSD(cell, 4 * nmol)
        d.set_precision(mode)
        d.set_graphs(graphs)
        d.set_orbitals(cell.random_orbitals(4 * nmol, seed=10 + i, decay=0.05))
        devs.append(d)
    return devs


@pytest.mark.parametrize("graphs,mode", [(False, 0), (True, 0), (False, 1), (True, 2)])
def test_batch_equals_sequential(gpu, graphs, mode):
    say("4 cells batched vs one after the other, graphs=%s, precision mode %d" % (graphs, mode))
    seq = make(False, mode)
    bat = make(graphs, mode)
    for call in range(4):
        ref = [d.run(5) for d in seq]
        got = cpsd_amd.run_batch(bat, 5)
        for (Er, der, dcr), (Eg, deg, dcg) in zip(ref, got):
            assert np.array_equal(Er, Eg), call
            assert der == deg and dcr == dcg
    for a, b in zip(seq, bat):
        x = a.get(c2=True, dn=True, hml=True)
        y = b.get(c2=True, dn=True, hml=True)
        for k in x:
            assert np.array_equal(x[k], y[k]), k
    for d in seq + bat:
        d.close()


def test_run_finish_needs_a_started_call(gpu):
    nmol, L, n = CELLS[0]
    cell = cpsd_amd.water_cell(nmol, L, n, seed=0)
    d = cpsd_amd.DeviceSD(cell, 4)
    d.set_orbitals(cell.random_orbitals(4, seed=1, decay=0.05))
    with pytest.raises(cpsd_amd.CpsdError):
        d.run_finish()
    d.run_start(2)
    with pytest.raises(cpsd_amd.CpsdError):
        d.run_start(2)
    d.run_finish()
    d.close()


@pytest.mark.parametrize("mode,learned", [(0, False), (1, True)])
def test_one_graph_batch_equals_sequential(gpu, mode, learned):
    """cpsd_batch: after the first (eager) call, every call of all cells is
    one captured graph replayed by a single launch; results equal the cells
    run one after the other, bit for bit, with uneven grids and orbital
    counts, and with the mixed mode and a preconditioner"""
    say("one-graph batch, mode %d, preconditioner %s" % (mode, learned))
    seq = make(False, mode)
    bat = make(False, mode)
    if learned:
        for d in seq + bat:
            d.set_preconditioner(np.full(d.cell.nfft3d, 2.0))
    b = cpsd_amd.Batch(bat)
    for call in range(5):
        ref = [d.run(5) for d in seq]
        got = b.run(5)
        for (Er, der, dcr), (Eg, deg, dcg) in zip(ref, got):
            assert np.array_equal(Er, Eg), call
            assert der == deg and dcr == dcg
    info = b.info()
    # call 1 eager; an odd inner alternates the orbital block between two
    # start states, each captured once and then replayed
    assert info == dict(graph_launches=4, captures=2, eager=1)
    b.close()
    for a, c in zip(seq, bat):
        x = a.get(c2=True, dn=True, hml=True)
        y = c.get(c2=True, dn=True, hml=True)
        for k in x:
            assert np.array_equal(x[k], y[k]), k
    for d in seq + bat:
        d.close()


def test_autotune_keeps_parity_and_is_remembered(gpu):
    """the autotuner's plan for a 128-orbital cell: the orbitals survive the
    measurement, the steps afterwards still match the oracle, and a second
    context of the same size class takes the remembered plan (no
    measurement) and runs the same bits"""
    import oracle_lib as O
    from test_gpu_parity import compare
    say("autotune")
    nfft, L, ne = 24, 12.0, 128
    cell = cpsd_amd.water_cell(2, L, nfft, seed=7)
    psi = cell.random_orbitals(ne, seed=3, decay=0.05)
    a = cpsd_amd.DeviceSD(cell, ne)
    a.set_orbitals(psi)
    plan = a.autotune(2)
    assert not plan["from_table"] and plan["splitk_target"] in (1024, 2048, 4096)
    st = O.State(O.Setup(1, L, nfft, cell.katm, cell.zv), ne, cell.lmmax, cell.vl, cell.vnl, cell.vnlnrm,
                 cell.r, psi, cell.amass)
    for _ in range(2):
        compare(a.run(3)[0], st.step(3)[0])
    assert np.max(np.abs(a.get(c2=True)["c2"] - st.c2)) < 1e-10
    b = cpsd_amd.DeviceSD(cell, ne)
    b.set_orbitals(psi)
    plan_b = b.autotune(2)
    assert plan_b["from_table"] and {k: v for k, v in plan_b.items() if k != "from_table"} == \
        {k: v for k, v in plan.items() if k != "from_table"}
    a.set_orbitals(psi)
    Ea = [a.run(3)[0] for _ in range(2)]
    Eb = [b.run(3)[0] for _ in range(2)]
    assert np.array_equal(np.array(Ea), np.array(Eb))
    a.close()
    b.close()


def test_batch64_one_graph_equals_single_cells_and_oracle(gpu):
    """the batched workload's 64 uneven cells (grids 24^3-64^3, 4-32
    orbitals; cpsd_amd.water_batch_specs(0)) in one graph per call, FP64:
    every cell's energies, deltas and orbitals equal the same cell run alone,
    bit for bit, and the three first 24^3 cells match the oracle (energies
    1e-9 Ha, c2 1e-10)"""
    import oracle_lib as O
    from test_gpu_parity import compare
    specs = cpsd_amd.water_batch_specs(0)
    assert len(specs) == 64 and len({(n, 4 * m) for m, L, n in specs}) >= 8
    say("batch-64: %d size classes" % len({(n, 4 * m) for m, L, n in specs}))
    cells = [cpsd_amd.water_cell(m, L, n, seed=i) for i, (m, L, n) in enumerate(specs)]
    psis = [c.random_orbitals(4 * m, seed=100 + i, decay=0.05) for i, (c, (m, L, n)) in
            enumerate(zip(cells, specs))]
    calls, inner = 3, 5
    alone = []
    for cell, psi, (m, L, n) in zip(cells, psis, specs):
        d = cpsd_amd.DeviceSD(cell, 4 * m)
        d.set_orbitals(psi)
        res = [d.run(inner) for _ in range(calls)]
        alone.append((res, d.get(c2=True)["c2"]))
        d.close()
    say("batch-64: single cells done")
    devs = []
    for cell, psi, (m, L, n) in zip(cells, psis, specs):
        d = cpsd_amd.DeviceSD(cell, 4 * m)
        d.set_orbitals(psi)
        devs.append(d)
    b = cpsd_amd.Batch(devs)
    got = [b.run(inner) for _ in range(calls)]
    info = b.info()
    say("batch-64: %s" % info)
    assert info["graph_launches"] >= 1 and info["eager"] == 1
    for i, d in enumerate(devs):
        for call in range(calls):
            Er, der, dcr = alone[i][0][call]
            Eg, deg, dcg = got[call][i]
            assert np.array_equal(Er, Eg), (i, call)
            assert der == deg and dcr == dcg, (i, call)
        assert np.array_equal(alone[i][1], d.get(c2=True)["c2"]), i
    b.close()
    small = [i for i, (m, L, n) in enumerate(specs) if n == 24][:3]
    assert len(small) == 3
    for i in small:
        m, L, n = specs[i]
        cell = cells[i]
        st = O.State(O.Setup(1, L, n, cell.katm, cell.zv), 4 * m, cell.lmmax, cell.vl, cell.vnl, cell.vnlnrm,
                     cell.r, psis[i], cell.amass)
        for call in range(calls):
            compare(got[call][i][0], st.step(inner)[0])
        assert np.max(np.abs(devs[i].get(c2=True)["c2"] - st.c2)) < 1e-10
    for d in devs:
        d.close()
