"""The identity behind the carried <c1|c1> of the GPU step, checked on the
oracle's own SD step (CPU).

A step's constrained orbitals are c2f = c2 + dte c1 lmd, where lmd[i + j ne]
is column-major (ccpsdv3/cpsd.c:482-488). So under gcdotc

    <c2f|c2f> = G22 + dte (T1 + T1^T) + dte^2 lmd^T W,

with T1 = G21 lmd, W = G11 lmd, G21[i][j] = <c2_i|c1_j>. This is how
k_gram_next_w and k_gram_next form the next step's <c1|c1>.

The check forms it from an oracle step's c1 and lmd, with c2 recovered as
c2f - dte c1 lmd, and compares it with <c2f|c2f> formed directly by the
oracle's gcdotc. A transposed lmd or a missing term fails it: the
alternatives are checked to differ.
"""
import numpy as np

import oracle_lib as O
from example_inputs import water_oracle_state


def gram(nfft, a, b):
    return np.array([[O.gcdotc(nfft, x, y) for y in b] for x in a])


def test_carried_overlap_identity():
    ne, dt, fmass = 4, 5.8, 400000.0
    dte = dt / np.sqrt(fmass)
    st = water_oracle_state()
    st.step(3, dt, fmass)          # away from the start
    st.step(1, dt, fmass)          # the step checked
    n = st.s.nfft
    n2 = st.s.n2ft3d
    c1 = st.c1.reshape(ne, n2)
    c2f = st.c2.reshape(ne, n2)
    lmd = st.lmd.reshape(ne, ne).T          # L[i][j] = lmd[i + j ne]
    c2 = c2f - dte * (lmd.T @ c1)           # c2f_j = c2_j + dte sum_i c1_i L[i][j]
    G11 = gram(n, c1, c1)
    G21 = gram(n, c2, c1)
    G22 = gram(n, c2, c2)
    direct = gram(n, c2f, c2f)
    W = G11 @ lmd
    T1 = G21 @ lmd
    carried = G22 + dte * (T1 + T1.T) + dte * dte * (lmd.T @ W)
    scale = np.max(np.abs(direct))
    assert np.max(np.abs(carried - direct)) < 1e-13 * scale
    # the constrained orbitals are orthonormal, so both are near the identity
    assert np.max(np.abs(direct - np.eye(ne))) < 1e-12
    # a dropped dte^2 term does not pass.  G21 and lmd are symmetric here
    # (<c2_i|c1_j> = <c1_i|c1_j> + dte <-H c1_i|c1_j> with H Hermitian, and the
    # Lagrange iteration keeps lmd symmetric), so a transposition of either
    # would not change the result and is not a mistake this can show
    err = np.max(np.abs(carried - direct))
    dropped = G22 + dte * (T1 + T1.T)
    assert np.max(np.abs(dropped - direct)) > 1e4 * max(err, 1e-300)
    assert np.max(np.abs(G21 - G21.T)) < 1e-15 and np.max(np.abs(lmd - lmd.T)) < 1e-15
