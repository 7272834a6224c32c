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
