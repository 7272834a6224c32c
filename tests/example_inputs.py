"""Shared builders for the reference's Example0 (C2 dimer) and the water
workloads, on top of the oracle (tests only)."""
import json
import os

import numpy as np

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def example0_expected():
    with open(os.path.join(GOLDEN, "example0", "expected.json")) as f:
        return json.load(f)


def water_expected():
    with open(os.path.join(GOLDEN, "h2o", "expected.json")) as f:
        return json.load(f)


def water_positions(dx=0.0, dy=0.0):
    """the recorded water geometry with both H moved by dx outward (along
    +-x) and by dy along y: the two symmetric directions in which the
    6-decimal rounding of the recorded H coordinates can move them"""
    exp = water_expected()
    r = np.array(exp["positions_bohr"], dtype=float)
    r[3] += dx; r[6] -= dx
    r[4] += dy; r[7] += dy
    return r


def rounding_window(energy_of, h=1e-5):
    """How far E(200) can move when each recorded H coordinate moves by up
    to half a unit of its last recorded decimal: sum over the 4 rounded
    coordinates of |dE/dq| * 5e-7, from forward differences of energy_of
    (a function of the positions).  O sits at the exact origin and the H
    z = 0 exactly, so they carry no rounding.  By the mirror symmetry
    x -> -x, dE/dx_H1 = -dE/dx_H2 and dE/dy_H1 = dE/dy_H2, so two symmetric
    displacements give all four derivatives."""
    exp = water_expected()
    half = 0.5 * 10.0 ** (-exp["positions_recorded_decimals"])
    e0 = energy_of(water_positions())
    gx = (energy_of(water_positions(dx=h)) - e0) / h / 2.0   # per H coordinate
    gy = (energy_of(water_positions(dy=h)) - e0) / h / 2.0
    return e0, 2.0 * half * (abs(gx) + abs(gy))


def read_control(path):
    t = open(path).read().split()
    return dict(ispin=int(t[0]), imove=int(t[1]), icube=int(t[2]), unit=float(t[3]),
                dt=float(t[4]), fmass=float(t[5]), inner=int(t[6]), outer=int(t[7]),
                tole=float(t[8]), tolc=float(t[9]), tolr=float(t[10]))


def read_ionin(path):
    """IONIN format of ccpsdv3/main.c:218-251"""
    lines = open(path).read().split("\n")
    nk = int(lines[0].split()[0])
    kinds = []
    for i in range(nk):
        q = lines[1 + i].split("'")
        kinds.append((q[1].strip(), int(q[2].split()[0])))
    toks = " ".join(lines[1 + nk:]).split()
    nion = int(toks[0])
    katm, r = [], []
    for i in range(nion):
        row = toks[1 + 7 * i: 1 + 7 * (i + 1)]
        katm.append(int(row[0]) - 1)
        r += [float(x) for x in row[1:4]]
    return kinds, np.array(katm, dtype=np.int32), np.array(r)


def example0_oracle_state():
    d = os.path.join(GOLDEN, "example0")
    ctl = read_control(os.path.join(d, "cpsdv3.d"))
    kinds, katm, r = read_ionin(os.path.join(d, "c2.ion00"))
    psp = O.read_psp(os.path.join(O.DATA, "psp", "C01.dat"))
    s = O.Setup(ctl["icube"], ctl["unit"], 32, katm, [psp["zv"]])
    vl, vnl, nrm = O.kbpp(psp, 2, s)
    c = O.makepsi(32, 4, O.read_psi0(os.path.join(O.DATA, "PSI0.DAT"), 4))
    st = O.State(s, 4, [4], vl, vnl, nrm, r, c, [psp["amass"] * 1822.89])
    return ctl, st


def water_oracle_state(nfft=32, L=20.0, ne=4, positions=None, katm=None, start="planewave", seed=1):
    """Oracle state of a water cell; start='planewave' uses PSI0.DAT."""
    pO = O.read_psp(os.path.join(O.DATA, "psp", "O00.dat"))
    pH = O.read_psp(os.path.join(O.DATA, "psp", "H00.dat"))
    if positions is None:
        positions = np.array([0, 0, 0, 1.431030, 1.108120, 0, -1.431030, 1.108120, 0.0])
        katm = [0, 1, 1]
    s = O.Setup(1, L, nfft, katm, [pO["zv"], pH["zv"]])
    vlO, vnlO, nO = O.kbpp(pO, 2, s)
    vlH, vnlH, nH = O.kbpp(pH, 1, s)
    vl = np.concatenate([vlO, vlH]); vnl = np.concatenate([vnlO, vnlH]); nrm = np.concatenate([nO, nH])
    c = O.makepsi(nfft, ne, O.read_psi0(os.path.join(O.DATA, "PSI0.DAT"), ne))
    return O.State(s, ne, [4, 1], vl, vnl, nrm, np.asarray(positions, float), c,
                   [16 * 1822.89, 1822.89])
