"""HIP-graph replay of whole calls (cpsd_dev_set_graphs) against uncaptured
calls, bit for bit.

Each case runs a graph context and an eager context from the same orbitals,
with the same arguments, call by call. Every call of the graph context is a
capture or a replay, and is compared with the eager call. The eager
context's own work and its synchronisations happen between the replays.

The cases cover:
- both parities of the orbital block (odd inner);
- a change of dt (a second graph), and a return to the first dt (a replay of
  the first);
- new orbitals in between;
- the mixed modes (FP32 transforms; with the BF16 KB projection, whose MFMA
  kernel runs in the 128-orbital case);
- 128 orbitals, where the overlap is carried from step to step.
"""
import sys

import numpy as np
import pytest

import cpsd_amd
from example_inputs import water_oracle_state
from test_gpu_parity import cell_from_state

pytestmark = pytest.mark.gpu


def say(msg):
    print("[gpu-graphs] " + msg, flush=True)
    sys.stdout.flush()


def pair(cell, ne, psi, mode):
    devs = []
    for graphs in (False, True):
        d = cpsd_amd.DeviceSD(cell, ne)
        d.set_precision(mode)
        d.set_graphs(graphs)
        d.set_orbitals(psi)
        devs.append(d)
    return devs


def same_call(devs, inner, dt=5.8, fmass=400000.0):
    Ee, de, ce = devs[0].run(inner, dt, fmass)
    Eg, dg, cg = devs[1].run(inner, dt, fmass)
    assert np.array_equal(Ee, Eg), (Ee, Eg)
    assert de == dg and ce == cg


def same_state(devs):
    a = devs[0].get(c1=True, c2=True, cpsi=True, dn=True, hml=True, lmd=True)
    b = devs[1].get(c1=True, c2=True, cpsi=True, dn=True, hml=True, lmd=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_graph_replays_water(gpu, mode):
    say("graph replay, water 32^3 ne=4, mode %d" % mode)
    st = water_oracle_state()
    cell = cell_from_state(st)
    devs = pair(cell, 4, st.c2, mode)
    same_call(devs, 5)            # the first call of a mode runs uncaptured
    for _ in range(4):            # capture both parities, then replay them
        same_call(devs, 5)
    same_state(devs)
    same_call(devs, 5, dt=4.0)    # another dt: another graph
    same_call(devs, 5, dt=4.0)
    same_call(devs, 5)            # back to the first one
    same_call(devs, 4)
    same_call(devs, 4)
    same_state(devs)
    for d in devs:                # new orbitals: the carried state restarts
        d.set_orbitals(st.c2)
    for _ in range(3):
        same_call(devs, 5)
    same_state(devs)
    for d in devs:
        d.close()


def test_graph_replays_carried_overlap(gpu):
    """128 orbitals at 24^3: the Gram with a carried <c1|c1> in the graphs,
    in all three precision modes"""
    say("graph replay, ne=128 at 24^3, modes 0-2")
    cell = cpsd_amd.water_cell(2, 12.0, 24, seed=7)
    psi = cell.random_orbitals(128, seed=8, decay=0.05)
    for mode in (0, 1, 2):
        devs = pair(cell, 128, psi, mode)
        for _ in range(5):
            same_call(devs, 3)
        same_state(devs)
        for d in devs:
            d.close()
